"""Build libfesom_tracer_amd.so (HIP for gfx950 + host C++) in-tree.

Run as ``python -m fesom2_tracer_dwarf_amd._build`` or through
``__graft_entry__.build()``. hipcc cross-compiles for gfx950 without a GPU.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
BUILD = PKG.parent / "_build"
LIBDIR = PKG / "lib"
LIBNAME = "libfesom_tracer_amd.so"
ALIAS = "libfesom_tracer_Fortran.so"
ROCM = Path(os.environ.get("ROCM_PATH", "/opt/rocm"))
ARCH = os.environ.get("FTA_OFFLOAD_ARCH", "gfx950")

# -ffp-contract=off keeps the FP64 path's operand order identical to the
# reference Fortran (no fused multiply-add), so parity is to rounding.
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # matrix-core results in ordinary VGPRs: the learned limiter's epilogue
    # reads every accumulator, and copies out of AGPRs would double its VALU work
    "-mllvm", "-amdgpu-mfma-vgpr-form",
    # device code assumes no NaN arithmetic: max/min then skip the IEEE-mode
    # quieting of each operand (one extra v_max per operand, a fifth of the
    # limiter's instructions); NaN results are still counted by bit tests
    "-Xarch_device", "-fno-honor-nans",
    "-Wno-unused-result",
]
CXX_FLAGS = ["-O2", "-fPIC", "-std=c++17", "-ffp-contract=off", "-D__HIP_PLATFORM_AMD__",
             f"-I{ROCM / 'include'}", "-Wall"]


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"command failed ({r.returncode}): {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")


def build(verbose: bool = False) -> Path:
    hipcc = shutil.which("hipcc") or str(ROCM / "bin" / "hipcc")
    cxx = shutil.which("g++") or "g++"
    BUILD.mkdir(exist_ok=True)
    LIBDIR.mkdir(exist_ok=True)
    jobs = []
    for src in sorted(CSRC.glob("*.hip")):
        jobs.append([hipcc, *HIP_FLAGS, "-c", str(src), "-o", str(BUILD / (src.stem + ".o"))])
    for src in sorted(CSRC.glob("*.cpp")):
        jobs.append([cxx, *CXX_FLAGS, "-c", str(src), "-o", str(BUILD / (src.stem + ".o"))])
    with ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        for j, _ in zip(jobs, ex.map(_run, jobs)):
            if verbose:
                print(" ".join(j))
    objs = [str(BUILD / (s.stem + ".o")) for s in sorted(CSRC.glob("*.hip")) + sorted(CSRC.glob("*.cpp"))]
    out = LIBDIR / LIBNAME
    tmp = LIBDIR / (LIBNAME + ".tmp")
    _run([hipcc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *objs,
          f"-L{ROCM / 'lib'}", "-lamdhip64", f"-Wl,-rpath,{ROCM / 'lib'}"])
    os.replace(tmp, out)
    # the file name the reference's Python wrapper loads (python/demo_tracer.py:27-37:
    # <its directory>/lib/libfesom_tracer_Fortran.so): the same library under that name
    alias = LIBDIR / ALIAS
    if alias.is_symlink() or alias.exists():
        alias.unlink()
    alias.symlink_to(LIBNAME)
    build_fortran(verbose)
    return out


FORTRAN_DIR = PKG / "fortran"
DRIVER = LIBDIR / "fesom_tracer_amd_driver"
# the FESOM-style host that tests the drop-in do_oce_adv_tra module, built
# against the minimal FESOM type modules of tests/fortran
TEST_FORTRAN = PKG.parent / "tests" / "fortran"
CHECKER = TEST_FORTRAN / "bin" / "check_do_oce_adv_tra"


def build_fortran(verbose: bool = False) -> Path | None:
    """The ISO_C_BINDING module, the Fortran analytic driver and (when the
    tests are present) the do_oce_adv_tra checker, with amdflang if it is
    installed; the programs link the library with $ORIGIN rpaths."""
    fc = shutil.which("amdflang") or (str(ROCM / "bin" / "amdflang") if (ROCM / "bin" / "amdflang").exists() else None)
    if not fc:
        return None
    mod = BUILD / "fesom_tracer_amd_f.o"
    cmds = [[fc, "-O2", "-c", str(FORTRAN_DIR / "fesom_tracer_amd.F90"), "-J", str(BUILD), "-o", str(mod)],
            [fc, "-O2", "-I", str(BUILD), str(FORTRAN_DIR / "analytic_driver.F90"), str(mod), f"-L{LIBDIR}",
             "-lfesom_tracer_amd", "-Wl,-rpath,$ORIGIN", "-o", str(DRIVER)]]
    if (TEST_FORTRAN / "fesom_types_min.F90").exists():
        tb = BUILD / "fortran_check"
        tb.mkdir(exist_ok=True)
        CHECKER.parent.mkdir(exist_ok=True)
        types_o, drv_o = tb / "fesom_types_min.o", tb / "oce_adv_tra_driver.o"
        cmds += [[fc, "-O2", "-c", str(TEST_FORTRAN / "fesom_types_min.F90"), "-J", str(tb), "-o", str(types_o)],
                 [fc, "-O2", "-c", str(FORTRAN_DIR / "oce_adv_tra_driver.F90"), "-I", str(tb), "-I", str(BUILD),
                  "-J", str(tb), "-o", str(drv_o)],
                 [fc, "-O2", "-I", str(tb), "-I", str(BUILD), str(TEST_FORTRAN / "check_do_oce_adv_tra.F90"),
                  str(types_o), str(drv_o),
                  str(mod), f"-L{LIBDIR}", "-lfesom_tracer_amd",
                  "-Wl,-rpath,$ORIGIN/../../../fesom2_tracer_dwarf_amd/lib", "-o", str(CHECKER)]]
    for c in cmds:
        if verbose:
            print(" ".join(c))
        _run(c)
    return DRIVER


if __name__ == "__main__":
    p = build(verbose="-v" in sys.argv)
    print(p)
