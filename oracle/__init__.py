"""CPU FP64 oracle for the tracer-advection step -- TEST INFRASTRUCTURE ONLY.

ctypes binding of oracle/libfesom_oracle.so (built from fesom_oracle.c by
oracle/Makefile) and of its OpenMP build libfesom_oracle_omp.so (the same
source; the edge-to-node scatters run as gathers in edge order, so its
results equal the serial build's bit for bit). Only tests, the smoke check
and the benchmark's CPU baseline use it, to compare against or to time; the
product never imports it.
"""
from __future__ import annotations

import ctypes
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "libfesom_oracle.so"
LIB_OMP = HERE / "libfesom_oracle_omp.so"

HOR = {"UPW1": 0, "MUSCL": 1, "MFCT": 2}
VER = {"UPW1": 0, "QR4C": 1, "PPM": 2, "CDIFF": 3}
LIM = {"NONE": 0, "FCT": 1}

_ip = ctypes.POINTER(ctypes.c_int)
_dp = ctypes.POINTER(ctypes.c_double)


class OracleMesh(ctypes.Structure):
    _fields_ = [("nl", ctypes.c_int), ("nod2D", ctypes.c_int), ("elem2D", ctypes.c_int),
                ("edge2D", ctypes.c_int), ("maxAdj", ctypes.c_int),
                ("edges", _ip), ("edge_tri", _ip), ("elem2D_nodes", _ip),
                ("nlevels", _ip), ("ulevels", _ip), ("nlevels_nod2D", _ip), ("ulevels_nod2D", _ip),
                ("nlevels_nod2D_min", _ip), ("ulevels_nod2D_max", _ip),
                ("nod_in_elem2D_num", _ip), ("nod_in_elem2D", _ip),
                ("coord_nod2D", _dp), ("edge_dxdy", _dp), ("edge_cross_dxdy", _dp),
                ("elem_cos", _dp), ("elem_area", _dp), ("gradient_sca", _dp),
                ("area", _dp), ("areasvol", _dp), ("hnode", _dp), ("hnode_new", _dp),
                ("helem", _dp), ("Z_3d_n", _dp), ("zbar_3d_n", _dp),
                ("edge_up_dn_tri", _ip), ("nboundary_lay", _ip)]


class OracleWork(ctypes.Structure):
    _fields_ = [(n, _dp) for n in ("fct_LO", "adv_flux_hor", "adv_flux_ver", "fct_ttf_max", "fct_ttf_min",
                                   "fct_plus", "fct_minus", "edge_up_dn_grad", "tr_xy", "dttf_h", "dttf_v",
                                   "del_ttf", "tvert_max", "tvert_min")]


MESH_INT = ("edges", "edge_tri", "elem2D_nodes", "nlevels", "ulevels", "nlevels_nod2D", "ulevels_nod2D",
            "nlevels_nod2D_min", "ulevels_nod2D_max", "nod_in_elem2D_num", "nod_in_elem2D",
            "edge_up_dn_tri", "nboundary_lay")
MESH_DBL = ("coord_nod2D", "edge_dxdy", "edge_cross_dxdy", "elem_cos", "elem_area", "gradient_sca", "area",
            "areasvol", "hnode", "hnode_new", "helem", "Z_3d_n", "zbar_3d_n")

_LIBS = {}


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB


def load(omp: bool = False):
    """The serial oracle, or its OpenMP build (omp=True)."""
    if omp not in _LIBS:
        path = LIB_OMP if omp else LIB
        if not path.exists():
            build()
        lib = ctypes.CDLL(str(path))
        P = ctypes.POINTER
        M, W = P(OracleMesh), P(OracleWork)
        d, i = ctypes.c_double, ctypes.c_int
        lib.oracle_step_tracer.argtypes = [M, d, _dp, _dp, _dp, _dp, _dp, i, i, i, d, d, W]
        lib.oracle_step_tracer.restype = i
        lib.oracle_step_tracer_ex.argtypes = [M, d, _dp, _dp, _dp, _dp, _dp, i, i, i, d, d, W, i]
        lib.oracle_step_tracer_ex.restype = i
        lib.oracle_do_oce_adv_tra.argtypes = [M, d, _dp, _dp, _dp, _dp, _dp, i, i, i, d, d, W]
        lib.oracle_do_oce_adv_tra.restype = i
        lib.oracle_tracer_gradient_elements.argtypes = [M, _dp, _dp]
        lib.oracle_fill_up_dn_grad.argtypes = [M, _dp, _dp]
        lib.oracle_find_up_downwind_triangles.argtypes = [M, d, d, _ip]
        lib.oracle_nboundary_lay.argtypes = [M, _ip]
        for nm in ("oracle_adv_tra_ver_upw1", "oracle_adv_tra_ver_cdiff"):
            getattr(lib, nm).argtypes = [M, _dp, _dp, _dp, i]
        lib.oracle_adv_tra_ver_qr4c.argtypes = [M, _dp, _dp, d, _dp, i]
        lib.oracle_adv_tra_vert_ppm.argtypes = [M, d, _dp, _dp, _dp, i]
        lib.oracle_adv_tra_hor_upw1.argtypes = [M, _dp, _dp, _dp, i]
        for nm in ("oracle_adv_tra_hor_muscl", "oracle_adv_tra_hor_mfct"):
            getattr(lib, nm).argtypes = [M, _dp, _dp, d, _dp, _dp, i]
        lib.oracle_tra_adv_fct.argtypes = [M, d, _dp, _dp, _dp, _dp, W]
        lib.oracle_tra_adv_flux2dtracer.argtypes = [M, d, _dp, _dp, _dp, _dp, i, _dp, _dp]
        lib.oracle_set_threads.argtypes = [i]
        lib.oracle_set_threads.restype = i
        _LIBS[omp] = lib
    return _LIBS[omp]


def _ptr(a, ct):
    return a.ctypes.data_as(ctypes.POINTER(ct))


class Oracle:
    """Holds one mesh (dict of arrays in the reference layout) and work arrays.
    threads=None runs the serial build; threads=n the OpenMP build on n threads
    (0: OpenMP's default), whose results are bitwise the same."""

    def __init__(self, mesh: dict, threads: int | None = None):
        self.lib = load(omp=threads is not None)
        self.threads = self.lib.oracle_set_threads(threads) if threads is not None else 1
        self.arrays = {}
        for k in MESH_INT:
            self.arrays[k] = np.ascontiguousarray(mesh[k], dtype=np.int32)
        for k in MESH_DBL:
            self.arrays[k] = np.ascontiguousarray(mesh[k], dtype=np.float64)
        nl, N, T, E, maxAdj = (int(mesh[k]) for k in ("nl", "nod2D", "elem2D", "edge2D", "maxAdj"))
        self.nl, self.N, self.T, self.E = nl, N, T, E
        if E < T:
            raise ValueError("FCT reuses edge_up_dn_grad(4,nl-1,edge2D) as element scratch: needs edge2D >= elem2D")
        m = OracleMesh(nl=nl, nod2D=N, elem2D=T, edge2D=E, maxAdj=maxAdj)
        for k in MESH_INT:
            setattr(m, k, _ptr(self.arrays[k], ctypes.c_int))
        for k in MESH_DBL:
            setattr(m, k, _ptr(self.arrays[k], ctypes.c_double))
        self.mesh = m
        L = nl - 1
        shapes = dict(fct_LO=L * N, adv_flux_hor=L * E, adv_flux_ver=nl * N, fct_ttf_max=L * N,
                      fct_ttf_min=L * N, fct_plus=L * N, fct_minus=L * N, edge_up_dn_grad=4 * L * E,
                      tr_xy=2 * L * T, dttf_h=L * N, dttf_v=L * N, del_ttf=L * N, tvert_max=L * N,
                      tvert_min=L * N)
        self.work_arrays = {k: np.zeros(v) for k, v in shapes.items()}
        self.work = OracleWork(**{k: _ptr(v, ctypes.c_double) for k, v in self.work_arrays.items()})

    def step(self, dt, vel, w, we, values, valuesAB=None, hor="MUSCL", ver="QR4C", lim="FCT", ph=1.0, pv=1.0,
             keep_del_ttf=False):
        """One analytic-driver step of one tracer; values (nl-1, nod2D) F-order is updated in place.
        keep_del_ttf: del_ttf carried over from this oracle's previous step (the reference C interface)."""
        assert values.flags.f_contiguous and values.dtype == np.float64
        ab = values if valuesAB is None else np.asfortranarray(valuesAB, dtype=np.float64)
        vel = np.ascontiguousarray(np.ravel(vel, order="F"))
        w = np.ascontiguousarray(np.ravel(w, order="F"))
        we = np.ascontiguousarray(np.ravel(we, order="F"))
        ab_copy = np.array(ab, order="F", copy=True)  # values is updated in place; AB must stay the old field
        st = self.lib.oracle_step_tracer_ex(ctypes.byref(self.mesh), dt, _ptr(vel, ctypes.c_double),
                                            _ptr(w, ctypes.c_double), _ptr(we, ctypes.c_double),
                                            _ptr(values, ctypes.c_double), _ptr(ab_copy, ctypes.c_double),
                                            HOR[hor], VER[ver], LIM[lim], ph, pv, ctypes.byref(self.work),
                                            1 if keep_del_ttf else 0)
        if st != 0:
            raise ValueError("oracle: unknown scheme")
        return values

    def tendencies(self, dt, vel, w, we, values, valuesAB=None, hor="MUSCL", ver="QR4C", lim="FCT", ph=1.0,
                   pv=1.0, dttf_h=None, dttf_v=None):
        """do_oce_adv_tra of one tracer (MUSCL/MFCT prerequisites from valuesAB first, as FESOM's
        caller computes them): returns dttf_h, dttf_v (nl-1, nod2D) after the routine added its
        tendencies into the given starting arrays (zeros when None); values is not changed."""
        L, N = self.nl - 1, self.N
        ab = values if valuesAB is None else valuesAB
        vals = np.asfortranarray(values, dtype=np.float64)
        ab = np.asfortranarray(ab, dtype=np.float64)
        vel = np.ascontiguousarray(np.ravel(vel, order="F"))
        w = np.ascontiguousarray(np.ravel(w, order="F"))
        we = np.ascontiguousarray(np.ravel(we, order="F"))
        wk = self.work_arrays
        wk["dttf_h"][:] = 0.0 if dttf_h is None else np.ravel(dttf_h, order="F")
        wk["dttf_v"][:] = 0.0 if dttf_v is None else np.ravel(dttf_v, order="F")
        if hor in ("MUSCL", "MFCT"):
            self.lib.oracle_tracer_gradient_elements(ctypes.byref(self.mesh), _ptr(ab, ctypes.c_double),
                                                     _ptr(wk["tr_xy"], ctypes.c_double))
            self.lib.oracle_fill_up_dn_grad(ctypes.byref(self.mesh), _ptr(wk["tr_xy"], ctypes.c_double),
                                            _ptr(wk["edge_up_dn_grad"], ctypes.c_double))
        st = self.lib.oracle_do_oce_adv_tra(ctypes.byref(self.mesh), dt, _ptr(vel, ctypes.c_double),
                                            _ptr(w, ctypes.c_double), _ptr(we, ctypes.c_double),
                                            _ptr(vals, ctypes.c_double), _ptr(ab, ctypes.c_double),
                                            HOR[hor], VER[ver], LIM[lim], ph, pv, ctypes.byref(self.work))
        if st != 0:
            raise ValueError("oracle: unknown scheme")
        return (wk["dttf_h"].reshape(N, L).T.copy(order="F"), wk["dttf_v"].reshape(N, L).T.copy(order="F"))

    def up_downwind(self, cyc_x_rad, cyc_y_rad):
        out = np.zeros(2 * self.E, dtype=np.int32)
        self.lib.oracle_find_up_downwind_triangles(ctypes.byref(self.mesh), cyc_x_rad, cyc_y_rad,
                                                   _ptr(out, ctypes.c_int))
        return out

    def nboundary_lay(self):
        out = np.zeros(self.N, dtype=np.int32)
        self.lib.oracle_nboundary_lay(ctypes.byref(self.mesh), _ptr(out, ctypes.c_int))
        return out
