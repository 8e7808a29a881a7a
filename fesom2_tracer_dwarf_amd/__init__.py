"""MI355X-native FESOM2 tracer-advection step (do_oce_adv_tra) behind the
reference's tracer_c_interface C ABI.

The compute path is libfesom_tracer_amd.so (HIP kernels for gfx950 plus a C
ABI); this package holds its ctypes binding, the Python wrapper class with
the reference's method names, and the synthetic workload definitions.
"""
from .tracer import FESOMTracerAdvection  # noqa: F401
from . import configs  # noqa: F401

__all__ = ["FESOMTracerAdvection", "configs"]
