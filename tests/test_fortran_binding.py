"""The ISO_C_BINDING layer (fesom2_tracer_dwarf_amd/fortran) against the C header.

CPU: every bind(C) name of the Fortran module is a function the header
declares and the library exports, and the reference's entry points bound
there keep their names. GPU: the Fortran analytic driver, run as its own
process, prints exactly the statistics the Python wrapper gets for the same
setup through the same C ABI.
"""
import re
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
MOD = ROOT / "fesom2_tracer_dwarf_amd" / "fortran" / "fesom_tracer_amd.F90"
DRIVER = ROOT / "fesom2_tracer_dwarf_amd" / "lib" / "fesom_tracer_amd_driver"


def _bound_names():
    return re.findall(r'bind\(C,\s*name="(\w+)"\)', MOD.read_text())


def test_fortran_binds_declared_functions():
    hdr = (ROOT / "include" / "fesom_tracer_amd.h").read_text()
    declared = set(re.findall(r"\b(tracer_\w+)\s*\(", hdr))
    names = _bound_names()
    assert len(names) >= 20
    missing = [n for n in names if n not in declared]
    assert not missing, f"bound but not declared: {missing}"
    for ref in ("tracer_init_mpi", "tracer_advect_step", "tracer_get_stats", "tracer_get_size",
                "tracer_get_values", "tracer_allocate_tracers", "tracer_set_values"):
        assert ref in names


def test_library_exports_fortran_names():
    from fesom2_tracer_dwarf_amd import _lib
    lib = _lib.load()
    for n in _bound_names():
        assert hasattr(lib, n), n


@pytest.mark.gpu
def test_fortran_driver_matches_python(api):
    if not DRIVER.exists():
        pytest.skip("Fortran driver not built (no amdflang)")
    print("[gpu] Fortran driver 32x24x16, 5 steps", flush=True)
    r = subprocess.run([str(DRIVER), "32", "24", "16", "5"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    lines = [l for l in r.stdout.splitlines() if l.startswith("step")]
    fort = np.array([[float(v) for v in l.split(":")[1].split()[4:]] for l in lines])
    s_line = [l for l in r.stdout.splitlines() if l.startswith("S active")][0]
    fort_s = np.array([float(v) for v in s_line.split()[-3:]])
    # the same workflow through the Python wrapper
    nx, ny, nl = 32, 24, 16
    Lx, Ly = 5.0e3 * nx, 5.0e3 * ny
    api.initialize_mpi(0)
    api.generate_mesh(nx, ny, nl, periodic=True, Lx=Lx, Ly=Ly, dz=10.0, stretch=0.0, lev_lo=3, lev_hi=nl, seed=1,
                      sfc=True)
    api.set_option("precision", 64)
    coord = api.mesh_array("coord")
    x, y = coord[0::2], coord[1::2]
    zn = np.arange(nl - 1) / (nl - 2)
    T = 15.0 + 5.0 * np.sin(2 * np.pi * x / Lx) * np.cos(2 * np.pi * y / Ly)
    T = np.asfortranarray(T[None, :] - 10.0 * zn[:, None])
    S = np.asfortranarray(34.0 + (0.5 + 0.5 * np.sin(2 * np.pi * x / Lx))[None, :] * (1.0 - zn)[:, None])
    api.allocate_tracers(2)
    api.set_values(1, T)
    api.set_values(2, S)
    for i in (1, 2):
        api.set_schemes(i, "MUSCL", "QR4C", "FCT")
    dt = api.make_velocity(0.1, 0.05, 0.2, 0.5, 0.3)
    py = []
    for _ in range(5):
        api.advect_steps(1, dt)
        st = api.get_statistics()
        py.append([st["min"], st["max"], st["sum"]])
    d = api.diagnostics(2)
    np.testing.assert_allclose(fort, np.array(py), rtol=1e-14, atol=0)
    np.testing.assert_allclose(fort_s, [d["active_min"], d["active_max"], d["mass"]], rtol=1e-14, atol=0)
