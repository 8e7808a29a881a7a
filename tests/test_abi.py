"""The C ABI library: it loads, exports every symbol include/fesom_tracer_amd.h
declares (the reference's 23 tracer_c_interface names among them), and its
host-side error paths return the reference's codes. No GPU compute here."""
import ctypes
import re
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]

# the names python/verify_c_interface.py and demo_tracer.py bind, plus the rest of
# lib/tracer_c_interface.F90's bind(C) functions
REFERENCE_NAMES = [
    "tracer_init_mpi", "tracer_finalize_mpi", "tracer_init", "tracer_advect_step", "tracer_get_stats",
    "tracer_get_size", "tracer_get_values", "tracer_finalize", "tracer_run_workflow", "tracer_cleanup",
    "tracer_set_mesh_dims", "tracer_set_levels", "tracer_set_thickness", "tracer_set_partition",
    "tracer_allocate_tracers", "tracer_set_values", "tracer_set_velocity_zero", "tracer_set_velocity",
    "tracer_init_complete", "tracer_load_mesh_partition", "tracer_load_mesh_from_files",
    "tracer_init_ale_arrays", "tracer_init_for_advection",
]


def header_symbols():
    text = (ROOT / "include" / "fesom_tracer_amd.h").read_text()
    return sorted(set(re.findall(r"^\s*(?:int|const char \*)\s*(tracer_\w+)\s*\(", text, re.M)))


def test_library_exports_every_declared_symbol():
    from fesom2_tracer_dwarf_amd import _lib
    lib = ctypes.CDLL(str(_lib.LIB_PATH))
    names = header_symbols()
    assert len(names) >= 23
    for name in names:
        assert hasattr(lib, name), name
    assert set(REFERENCE_NAMES) <= set(names)
    assert set(_lib.SIGNATURES) == set(names)


def test_error_codes_without_initialisation(api):
    lib = api.lib
    assert lib.tracer_cleanup() == 0
    assert lib.tracer_advect_step(1.0) == -1
    tmin, tmax, tsum = ctypes.c_double(1), ctypes.c_double(1), ctypes.c_double(1)
    assert lib.tracer_get_stats(ctypes.byref(tmin), ctypes.byref(tmax), ctypes.byref(tsum)) == -1
    assert (tmin.value, tmax.value, tsum.value) == (0.0, 0.0, 0.0)
    nz, nn = ctypes.c_int(5), ctypes.c_int(5)
    assert lib.tracer_get_size(ctypes.byref(nz), ctypes.byref(nn)) == -1 and nz.value == 0
    assert lib.tracer_init(b"/nonexistent/restart") == -1
    assert lib.tracer_load_mesh_partition(b"/nonexistent/restart") == -1
    assert lib.tracer_run_workflow(b"/nonexistent/restart", 2, 1.0) == -1
    assert lib.tracer_init_mpi(0) == 0 and lib.tracer_finalize_mpi() == 0
    assert "restart" in api.lib.tracer_last_error().decode()


def test_host_side_checks_after_mesh(api):
    d = api.generate_mesh(6, 5, 5, periodic=True, Lx=6e4, Ly=5e4, lev_lo=5, lev_hi=5)
    api.allocate_tracers(2)
    v = np.zeros((d["nl"] - 1, d["nod2D"]), order="F")
    assert api.lib.tracer_set_values(3, v, d["nl"], d["nod2D"]) == -1       # invalid id
    assert api.lib.tracer_set_values(0, v, d["nl"], d["nod2D"]) == -1
    assert api.lib.tracer_set_values(1, v, d["nl"], d["nod2D"]) == 0
    with pytest.raises(RuntimeError, match="Unknown horizontal"):
        api.set_schemes(1, "WENO", "QR4C", "FCT")
    with pytest.raises(RuntimeError, match="Unknown vertical"):
        api.set_schemes(1, "MUSCL", "QR8", "FCT")
    with pytest.raises(RuntimeError, match="Unknown limiter"):
        api.set_schemes(1, "MUSCL", "QR4C", "ZALESAK")
    api.set_schemes(2, "MFCT", "PPM", "FCT")
    with pytest.raises(RuntimeError):
        api.set_option("precision", 8)
    with pytest.raises(RuntimeError):
        api.set_option("no_such_option", 1)
    nz, nn = ctypes.c_int(), ctypes.c_int()
    assert api.lib.tracer_get_size(ctypes.byref(nz), ctypes.byref(nn)) == 0
    assert (nz.value, nn.value) == (d["nl"] - 1, d["nod2D"])
    assert api.lib.tracer_set_partition(d["nod2D"], 3) == -1                # one partition only


def test_qr4c_needs_two_layers(api):
    api.generate_mesh(5, 5, 4, periodic=True, Lx=5e4, Ly=5e4, lev_lo=2, lev_hi=2)
    api.allocate_tracers(1)
    with pytest.raises(RuntimeError, match="two layers"):
        api.set_schemes(1, "MUSCL", "QR4C", "FCT")
    api.set_schemes(1, "MUSCL", "CDIFF", "FCT")


def test_wrapper_keeps_reference_method_names():
    from fesom2_tracer_dwarf_amd import FESOMTracerAdvection
    for name in ("initialize_mpi", "initialize", "advect_step", "get_tracer_values", "get_statistics",
                 "finalize", "finalize_mpi"):
        assert callable(getattr(FESOMTracerAdvection, name))


def test_velocity_buffer_is_read_as_element_velocity(api):
    """tracer_set_velocity's (nl-1, edge2D, 2) buffer is read, as in the
    reference, as do_oce_adv_tra's vel(2, nl-1, elem2D): its first
    2*(nl-1)*elem2D values, component fastest."""
    d = api.generate_mesh(6, 5, 5, periodic=True, Lx=6e4, Ly=5e4, lev_lo=5, lev_hi=5)
    L, T, E, N = d["nl"] - 1, d["elem2D"], d["edge2D"], d["nod2D"]
    buf = np.asfortranarray(np.arange(L * E * 2, dtype=np.float64).reshape(L, E, 2, order="F"))
    w = np.asfortranarray(np.ones((d["nl"], N)))
    assert api.lib.tracer_set_velocity(buf, w, d["nl"], N, E) == 0
    uv = api.mesh_array("uv")
    assert np.array_equal(uv, np.ravel(buf, order="F")[: 2 * L * T])
    assert np.all(api.mesh_array("w_e") == 0.0) and np.all(api.mesh_array("w") == 1.0)
