"""CPU checks of the drop-in boundary (no compute calls):

* python/demo_tracer.py:27-37 loads <its directory>/lib/libfesom_tracer_Fortran.so:
  the build provides that name as the same library, exporting the
  reference's 23 entry points;
* the Fortran drop-in module keeps do_oce_adv_tra's name and argument list
  (lib/oce_adv_tra_driver.F90:23) and compiled against FESOM-layout derived
  types, with a FESOM-style host program linked to it;
* the new entry points are declared and exported.
"""
import ctypes
import os
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
LIB = ROOT / "fesom2_tracer_dwarf_amd" / "lib"
DRV = ROOT / "fesom2_tracer_dwarf_amd" / "fortran" / "oce_adv_tra_driver.F90"


def test_reference_wrapper_library_name():
    from fesom2_tracer_dwarf_amd import _lib
    alias = LIB / "libfesom_tracer_Fortran.so"
    assert alias.exists() and os.path.samefile(alias, LIB / "libfesom_tracer_amd.so")
    lib = ctypes.CDLL(str(alias), mode=ctypes.RTLD_GLOBAL)
    for name in _lib.REFERENCE_SYMBOLS:
        assert hasattr(lib, name), name


def test_fortran_drop_in_keeps_the_signature():
    src = DRV.read_text()
    assert re.search(r"^module oce_adv_tra_driver_module", src, re.M)
    m = re.search(r"subroutine do_oce_adv_tra\(([^)]*)\)", src)
    assert m and [a.strip() for a in m.group(1).split(",")] == [
        "dt", "vel", "w", "wi", "we", "tr_num", "dynamics", "tracers", "partit", "mesh"]
    for mod in ("MOD_MESH", "MOD_TRACER", "MOD_PARTIT", "MOD_DYN"):
        assert re.search(rf"^\s*use {mod}\b", src, re.M | re.I)
    hdr = (ROOT / "include" / "fesom_tracer_amd.h").read_text()
    for name in re.findall(r"\b(tracer_\w+)\(", re.sub(r"!.*", "", src)):
        assert re.search(rf"\b{name}\s*\(", hdr), f"{name} used by the module but not declared"


def test_fortran_host_is_built():
    import shutil
    if not (shutil.which("amdflang") or Path("/opt/rocm/bin/amdflang").exists()):
        pytest.skip("no amdflang")
    assert (ROOT / "tests" / "fortran" / "bin" / "check_do_oce_adv_tra").exists()


def test_new_entry_points_exported():
    from fesom2_tracer_dwarf_amd import _lib
    lib = _lib.load()
    for name in ("tracer_do_oce_adv_tra", "tracer_hbm_copy_rate", "tracer_batch_get_mesh_array"):
        assert hasattr(lib, name)
