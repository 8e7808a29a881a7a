"""Batched form: several independent periodic meshes and their tracers,
advanced by one hipGraph launch per step (C ABI tracer_batch_*)."""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib


class TracerBatch:
    def __init__(self):
        self.lib = _lib.load()
        _lib.check(self.lib.tracer_batch_create(), "tracer_batch_create", self.lib)
        self.meshes = []

    def add_periodic_mesh(self, nx, ny, nl, Lx, Ly, dz=10.0, lev_lo=3, lev_hi=None, seed=0, ntracers=32,
                          precision=32, schemes=("MUSCL", "QR4C", "FCT"), cfl=0.3):
        dt = ctypes.c_double()
        i = self.lib.tracer_batch_add_periodic_mesh(nx, ny, nl, Lx, Ly, dz, lev_lo, lev_hi or nl, seed, ntracers,
                                                    precision, *(s.encode() for s in schemes), cfl,
                                                    ctypes.byref(dt))
        if i < 0:
            raise RuntimeError(f"tracer_batch_add_periodic_mesh failed: {_lib.last_error(self.lib)}")
        self.meshes.append(i)
        return i, dt.value

    def info(self, mesh):
        nl, n, ntr, c, dt = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_longlong(), ctypes.c_double()
        _lib.check(self.lib.tracer_batch_mesh_info(mesh, ctypes.byref(nl), ctypes.byref(n), ctypes.byref(ntr),
                                                   ctypes.byref(c), ctypes.byref(dt)), "tracer_batch_mesh_info",
                   self.lib)
        return dict(nl=nl.value, nod2D=n.value, ntracers=ntr.value, active_cells=c.value, dt=dt.value)

    def run(self, nsteps, use_graph=True):
        _lib.check(self.lib.tracer_batch_run(int(nsteps), 1 if use_graph else 0), "tracer_batch_run", self.lib)

    def autotune(self, reps=3):
        """Each mesh keeps its fastest launch schedule; returns [(waves, tracers per launch)] per mesh."""
        out = np.zeros(2 * max(len(self.meshes), 1))
        n = self.lib.tracer_batch_autotune(int(reps), out, len(self.meshes))
        if n < 0:
            raise RuntimeError(f"tracer_batch_autotune failed: {_lib.last_error(self.lib)}")
        return [(int(out[2 * i]), int(out[2 * i + 1])) for i in range(n)]

    def set_schedule(self, waves, max_nt):
        _lib.check(self.lib.tracer_batch_set_schedule(int(waves), int(max_nt)), "tracer_batch_set_schedule",
                   self.lib)

    def time(self, nsteps):
        ms = ctypes.c_double()
        _lib.check(self.lib.tracer_batch_time(int(nsteps), ctypes.byref(ms)), "tracer_batch_time", self.lib)
        return ms.value

    def values(self, mesh, tracer_id):
        d = self.info(mesh)
        out = np.zeros((d["nl"] - 1, d["nod2D"]), order="F")
        _lib.check(self.lib.tracer_batch_get_values(mesh, tracer_id, out), "tracer_batch_get_values", self.lib)
        return out

    def diagnostics(self, mesh, tracer_id):
        out = np.zeros(7)
        _lib.check(self.lib.tracer_batch_diagnostics(mesh, tracer_id, out), "tracer_batch_diagnostics", self.lib)
        return dict(zip(("min", "max", "sum", "active_min", "active_max", "mass", "nan_count"), out.tolist()))

    def mesh_array(self, mesh, name):
        """A named array of a batch mesh (names as FESOMTracerAdvection.mesh_array, plus uv, w, w_e)."""
        n, kind = ctypes.c_longlong(), ctypes.c_int()
        _lib.check(self.lib.tracer_batch_get_mesh_array(mesh, name.encode(), None, 0, ctypes.byref(n),
                                                        ctypes.byref(kind)), "tracer_batch_get_mesh_array", self.lib)
        buf = np.zeros(n.value, dtype=np.float64 if kind.value else np.int32)
        _lib.check(self.lib.tracer_batch_get_mesh_array(mesh, name.encode(), buf.ctypes.data, n.value, None, None),
                   "tracer_batch_get_mesh_array", self.lib)
        return buf

    def dims(self, mesh):
        """Dimensions of a batch mesh, as FESOMTracerAdvection.dims."""
        d = self.info(mesh)
        return dict(nl=d["nl"], nod2D=d["nod2D"], elem2D=int(self.mesh_array(mesh, "nlevels").size),
                    edge2D=int(self.mesh_array(mesh, "edges").size // 2),
                    maxAdj=int(self.mesh_array(mesh, "nod_in_elem2D").size // d["nod2D"]))

    def close(self):
        self.lib.tracer_batch_destroy()
