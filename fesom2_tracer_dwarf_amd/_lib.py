"""ctypes binding of libfesom_tracer_amd.so (C ABI in include/fesom_tracer_amd.h).

There is no fallback: if the library is missing or does not load, importing
code gets a RuntimeError naming the build command.
"""
from __future__ import annotations

import ctypes
import os
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "lib" / "libfesom_tracer_amd.so"

c_int, c_double, c_char_p, c_void_p = ctypes.c_int, ctypes.c_double, ctypes.c_char_p, ctypes.c_void_p
c_ll, c_ull = ctypes.c_longlong, ctypes.c_ulonglong
P_int, P_double, P_ll = ctypes.POINTER(c_int), ctypes.POINTER(c_double), ctypes.POINTER(c_ll)
F64 = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
F64F = np.ctypeslib.ndpointer(dtype=np.float64, flags="F_CONTIGUOUS")
I32 = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
F32 = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
c_float, c_uint, c_longlong, POINTER = ctypes.c_float, ctypes.c_uint, ctypes.c_longlong, ctypes.POINTER

# name -> (restype, argtypes); the 23 reference symbols first
SIGNATURES = {
    "tracer_init_mpi": (c_int, [c_int]),
    "tracer_finalize_mpi": (c_int, []),
    "tracer_init": (c_int, [c_char_p]),
    "tracer_advect_step": (c_int, [c_double]),
    "tracer_get_stats": (c_int, [P_double, P_double, P_double]),
    "tracer_get_size": (c_int, [P_int, P_int]),
    "tracer_get_values": (c_int, [F64F, c_int, c_int]),
    "tracer_finalize": (c_int, [c_char_p]),
    "tracer_run_workflow": (c_int, [c_char_p, c_int, c_double]),
    "tracer_cleanup": (c_int, []),
    "tracer_set_mesh_dims": (c_int, [c_int, c_int, c_int, c_int]),
    "tracer_set_levels": (c_int, [I32, I32, c_int]),
    "tracer_set_thickness": (c_int, [F64F, c_int, c_int]),
    "tracer_set_partition": (c_int, [c_int, c_int]),
    "tracer_allocate_tracers": (c_int, [c_int, c_int, c_int, c_int]),
    "tracer_set_values": (c_int, [c_int, F64F, c_int, c_int]),
    "tracer_set_velocity_zero": (c_int, [c_int, c_int, c_int]),
    "tracer_set_velocity": (c_int, [F64F, F64F, c_int, c_int, c_int]),
    "tracer_init_complete": (c_int, []),
    "tracer_load_mesh_partition": (c_int, [c_char_p]),
    "tracer_load_mesh_from_files": (c_int, [c_char_p]),
    "tracer_write_mesh": (c_int, [c_char_p]),
    "tracer_output_open": (c_int, [c_char_p, c_int]),
    "tracer_output_step": (c_int, []),
    "tracer_output_close": (c_int, []),
    "tracer_init_ale_arrays": (c_int, []),
    "tracer_init_for_advection": (c_int, []),
    # extensions
    "tracer_generate_mesh": (c_int, [c_int, c_int, c_int, c_double, c_double, c_int, c_double, c_double,
                                     c_int, c_int, c_ull, c_int]),
    "tracer_set_schemes": (c_int, [c_int, c_char_p, c_char_p, c_char_p, c_double, c_double]),
    "tracer_set_velocity_elem": (c_int, [F64F, F64F, F64F, c_int, c_int, c_int]),
    "tracer_make_velocity": (c_int, [c_double, c_double, c_double, c_double, c_double, P_double]),
    "tracer_set_values_ab": (c_int, [c_int, F64F, c_int, c_int]),
    "tracer_get_values_id": (c_int, [c_int, F64F, c_int, c_int]),
    "tracer_set_option": (c_int, [c_char_p, c_int]),
    "tracer_advect_steps": (c_int, [c_int, c_double]),
    "tracer_do_oce_adv_tra": (c_int, [c_int, c_double, F64F, F64F, c_int, c_int]),
    "tracer_do_oce_adv_tra_noalloc": (c_int, [c_int, c_double, c_void_p, c_void_p, c_int, c_int]),
    "tracer_get_diagnostics": (c_int, [c_int, F64]),
    "tracer_mesh_array_info": (c_int, [c_char_p, P_ll, P_int]),
    "tracer_get_mesh_array": (c_int, [c_char_p, c_void_p, c_ll]),
    "tracer_mesh_begin": (c_int, [c_int, c_int, c_int, c_int, c_int]),
    "tracer_restart_file_copy": (c_int, [c_char_p, c_char_p, c_char_p]),
    "tracer_set_mesh_array": (c_int, [c_char_p, c_void_p, c_ll]),
    "tracer_mesh_end": (c_int, []),
    "tracer_get_work_array": (c_int, [c_int, c_char_p, F64, c_ll]),
    "tracer_time_steps": (c_int, [c_int, c_double, P_double]),
    "tracer_time_kernels": (c_int, [c_double, c_int, c_char_p, c_int, F64, c_int]),
    "tracer_precision_error": (c_int, [c_int, c_int, c_double, F64]),
    "tracer_mlp_param_count": (c_int, []),
    "tracer_mlp_set_params": (c_int, [F32, c_int]),
    "tracer_mlp_get_params": (c_int, [F32, c_int]),
    "tracer_learn_collect": (c_int, [c_int, c_double, c_longlong, POINTER(c_longlong)]),
    "tracer_learn_clear": (c_int, []),
    "tracer_learn_train": (c_int, [c_int, c_longlong, c_double, c_uint, F64]),
    "tracer_mlp_forward": (c_int, [F32, c_longlong, F32]),
    "tracer_mlp_gradient": (c_int, [F32, F32, c_longlong, c_float, F32, POINTER(c_double)]),
    "tracer_learn_fallback": (c_int, [c_int, POINTER(c_double)]),
    "tracer_autotune": (c_int, [c_double, c_int, F64, c_int]),
    "tracer_hbm_copy_rate": (c_int, [c_longlong, c_int, POINTER(c_double), POINTER(c_int)]),
    "tracer_batch_create": (c_int, []),
    "tracer_batch_destroy": (c_int, []),
    "tracer_batch_add_periodic_mesh": (c_int, [c_int, c_int, c_int, c_double, c_double, c_double, c_int, c_int,
                                               c_ull, c_int, c_int, c_char_p, c_char_p, c_char_p, c_double,
                                               POINTER(c_double)]),
    "tracer_batch_run": (c_int, [c_int, c_int]),
    "tracer_batch_time": (c_int, [c_int, POINTER(c_double)]),
    "tracer_batch_autotune": (c_int, [c_int, F64, c_int]),
    "tracer_batch_set_schedule": (c_int, [c_int, c_int]),
    "tracer_batch_mesh_info": (c_int, [c_int, P_int, P_int, P_int, P_ll, P_double]),
    "tracer_batch_get_values": (c_int, [c_int, c_int, F64F]),
    "tracer_batch_diagnostics": (c_int, [c_int, c_int, F64]),
    "tracer_batch_get_mesh_array": (c_int, [c_int, c_char_p, c_void_p, c_ll, P_ll, P_int]),
    "tracer_last_error": (c_char_p, []),
}
REFERENCE_SYMBOLS = list(SIGNATURES)[:23]

_LIB = None


def load(path: str | os.PathLike | None = None) -> ctypes.CDLL:
    """Load the library once (RTLD_GLOBAL, as the reference wrapper does)."""
    global _LIB
    if _LIB is not None and path is None:
        return _LIB
    p = Path(path) if path else LIB_PATH
    if not p.exists():
        raise RuntimeError(f"{p} is missing: build it with `python -m fesom2_tracer_dwarf_amd._build` "
                           "(or __graft_entry__.build())")
    lib = ctypes.CDLL(str(p), mode=ctypes.RTLD_GLOBAL)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _LIB = lib
    return lib


def last_error(lib=None) -> str:
    lib = lib or load()
    s = lib.tracer_last_error()
    return s.decode() if s else ""


def check(status: int, what: str, lib=None) -> None:
    if status != 0:
        raise RuntimeError(f"{what} failed ({status}): {last_error(lib)}")
