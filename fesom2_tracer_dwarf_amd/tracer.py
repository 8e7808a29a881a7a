"""Python interface of the tracer-advection library.

``FESOMTracerAdvection`` keeps the method names, arguments and return values
of the reference's ctypes wrapper class (python/demo_tracer.py:40-208); the
extra methods below it drive the parts of the step the reference wrapper
cannot reach (mesh generation, schemes, element velocity, batched steps).
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib


MESH_INT_ARRAYS = frozenset({
    "edges", "edge_tri", "elem2D_nodes", "nlevels", "ulevels", "nlevels_nod2D", "ulevels_nod2D", "nlevels_nod2D_min",
    "ulevels_nod2D_max", "nod_in_elem2D_num", "nod_in_elem2D", "edge_up_dn_tri", "nboundary_lay"})


class FESOMTracerAdvection:
    """Handle on the library's single instance (python/demo_tracer.py:40)."""

    def __init__(self, lib_path=None):
        self.lib = _lib.load(lib_path)

    # ---- reference methods (python/demo_tracer.py:101-208) -------------
    def initialize_mpi(self, mpi_comm=0):
        return self.lib.tracer_init_mpi(mpi_comm)

    def initialize(self, restart_path="./fesom_bin_restart"):
        return self.lib.tracer_init(restart_path.encode())

    def advect_step(self, dt=1.0e-3):
        return self.lib.tracer_advect_step(dt)

    def get_tracer_values(self, tracer_id=1):
        nz, nn = ctypes.c_int(), ctypes.c_int()
        ret = self.lib.tracer_get_size(ctypes.byref(nz), ctypes.byref(nn))
        if ret != 0:
            raise RuntimeError(f"tracer_get_size failed with code {ret}")
        values = np.zeros((nz.value, nn.value), dtype=np.float64, order="F")
        ret = self.lib.tracer_get_values_id(tracer_id, values, nz, nn)
        if ret != 0:
            raise RuntimeError(f"tracer_get_values failed with code {ret}")
        return values

    def get_statistics(self):
        tmin, tmax, tsum = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        ret = self.lib.tracer_get_stats(ctypes.byref(tmin), ctypes.byref(tmax), ctypes.byref(tsum))
        if ret != 0:
            raise RuntimeError(f"tracer_get_stats failed with code {ret}")
        return {"min": tmin.value, "max": tmax.value, "sum": tsum.value}

    def finalize(self, output_path="./fesom_bin_restart"):
        return self.lib.tracer_finalize(output_path.encode())

    def finalize_mpi(self):
        return self.lib.tracer_finalize_mpi()

    # ---- extensions -----------------------------------------------------
    def generate_mesh(self, nx, ny, nl, *, periodic=True, Lx=1.0e5, Ly=1.0e5, dz=10.0, stretch=0.0,
                      lev_lo=None, lev_hi=None, seed=0, sfc=True):
        lev_lo = nl if lev_lo is None else lev_lo
        lev_hi = nl if lev_hi is None else lev_hi
        _lib.check(self.lib.tracer_generate_mesh(nx, ny, int(periodic), Lx, Ly, nl, dz, stretch, lev_lo, lev_hi,
                                                 seed, int(sfc)), "tracer_generate_mesh", self.lib)
        return self.dims()

    def load_mesh_partition(self, restart_path):
        """Mesh, partition and velocity from derived-type restarts under restart_path/np1 (no tracers)."""
        _lib.check(self.lib.tracer_load_mesh_partition(str(restart_path).encode()), "tracer_load_mesh_partition",
                   self.lib)
        return self.dims()

    def run_workflow(self, restart_path, nsteps, dt):
        """tracer_init from restart_path, nsteps steps of tracer 1, tracer_finalize into restart_path."""
        return self.lib.tracer_run_workflow(str(restart_path).encode(), nsteps, dt)

    def load_mesh_from_files(self, mesh_path):
        """FESOM mesh files (nod2d.out, elem2d.out, aux3d.out, elvls.out,
        nlvls.out, edges.out, edge_tri.out, edgenum.out) from a directory."""
        _lib.check(self.lib.tracer_load_mesh_from_files(str(mesh_path).encode()), "tracer_load_mesh_from_files",
                   self.lib)
        return self.dims()

    def write_mesh(self, output_dir):
        """mesh_info.txt and the binary mesh files the reference's plotting scripts read."""
        _lib.check(self.lib.tracer_write_mesh(str(output_dir).encode()), "tracer_write_mesh", self.lib)

    def output_open(self, output_dir, nsteps):
        _lib.check(self.lib.tracer_output_open(str(output_dir).encode(), int(nsteps)), "tracer_output_open",
                   self.lib)

    def output_step(self):
        _lib.check(self.lib.tracer_output_step(), "tracer_output_step", self.lib)

    def output_close(self):
        _lib.check(self.lib.tracer_output_close(), "tracer_output_close", self.lib)

    def dims(self):
        d = self.mesh_array("dims")
        return dict(nl=int(d[0]), nod2D=int(d[1]), elem2D=int(d[2]), edge2D=int(d[3]), maxAdj=int(d[4]))

    def mesh_array(self, name):
        cnt, kind = ctypes.c_longlong(), ctypes.c_int()
        _lib.check(self.lib.tracer_mesh_array_info(name.encode(), ctypes.byref(cnt), ctypes.byref(kind)),
                   f"mesh array {name}", self.lib)
        out = np.empty(cnt.value, dtype=np.float64 if kind.value else np.int32)
        _lib.check(self.lib.tracer_get_mesh_array(name.encode(), out.ctypes.data, cnt.value),
                   f"mesh array {name}", self.lib)
        return out

    def set_mesh(self, nl, nod2D, elem2D, edge2D, maxAdj, arrays):
        """A caller-supplied mesh: ``arrays`` maps the names of :meth:`mesh_array` to arrays in its layout
        (0-based indices, -1 for no element). Index arrays go in as int32, the rest as float64."""
        _lib.check(self.lib.tracer_mesh_begin(nl, nod2D, elem2D, edge2D, maxAdj), "tracer_mesh_begin", self.lib)
        for name, a in arrays.items():
            a = np.ascontiguousarray(np.ravel(a, order="A"), dtype=np.int32 if name in MESH_INT_ARRAYS else np.float64)
            _lib.check(self.lib.tracer_set_mesh_array(name.encode(), a.ctypes.data, a.size), f"mesh array {name}",
                       self.lib)
        _lib.check(self.lib.tracer_mesh_end(), "tracer_mesh_end", self.lib)
        return self.dims()

    def allocate_tracers(self, num_tracers, AB_order=1):
        self.num_tracers = int(num_tracers)
        d = self.dims()
        _lib.check(self.lib.tracer_allocate_tracers(num_tracers, d["nl"], d["nod2D"], AB_order),
                   "tracer_allocate_tracers", self.lib)

    def set_values(self, tracer_id, values):
        v = np.asfortranarray(values, dtype=np.float64)
        _lib.check(self.lib.tracer_set_values(tracer_id, v, v.shape[0] + 1, v.shape[1]), "tracer_set_values",
                   self.lib)

    def set_values_ab(self, tracer_id, values):
        v = np.asfortranarray(values, dtype=np.float64)
        _lib.check(self.lib.tracer_set_values_ab(tracer_id, v, v.shape[0] + 1, v.shape[1]),
                   "tracer_set_values_ab", self.lib)

    def set_schemes(self, tracer_id, hor="MUSCL", ver="QR4C", lim="FCT", ph=1.0, pv=1.0):
        _lib.check(self.lib.tracer_set_schemes(tracer_id, hor.encode(), ver.encode(), lim.encode(), ph, pv),
                   "tracer_set_schemes", self.lib)

    def make_velocity(self, u_bg=0.1, v_bg=0.05, cell_amp=0.2, shear=0.5, cfl=0.3):
        dt = ctypes.c_double()
        _lib.check(self.lib.tracer_make_velocity(u_bg, v_bg, cell_amp, shear, cfl, ctypes.byref(dt)),
                   "tracer_make_velocity", self.lib)
        return dt.value

    def set_thickness(self, hnode_new):
        """hnode_new, (nl-1, nod2D) Fortran order (tracer_set_thickness: hnode stays)."""
        d = self.dims()
        _lib.check(self.lib.tracer_set_thickness(np.asfortranarray(hnode_new, dtype=np.float64), d["nl"], d["nod2D"]),
                   "tracer_set_thickness", self.lib)

    def set_velocity_elem(self, uv, w, w_e):
        d = self.dims()
        _lib.check(self.lib.tracer_set_velocity_elem(np.asfortranarray(uv, dtype=np.float64),
                                                     np.asfortranarray(w, dtype=np.float64),
                                                     np.asfortranarray(w_e, dtype=np.float64),
                                                     d["nl"], d["nod2D"], d["elem2D"]),
                   "tracer_set_velocity_elem", self.lib)

    def set_option(self, name, value):
        _lib.check(self.lib.tracer_set_option(name.encode(), int(value)), f"option {name}", self.lib)

    def advect_steps(self, nsteps, dt):
        _lib.check(self.lib.tracer_advect_steps(nsteps, dt), "tracer_advect_steps", self.lib)

    def do_oce_adv_tra(self, tracer_id, dt, dttf_h=None, dttf_v=None):
        """One do_oce_adv_tra of a tracer: its tendencies added into dttf_h /
        dttf_v ((nl-1, nod2D) F-order float64, zeros when not given), which are
        returned; the tracer is unchanged."""
        d = self.dims()
        shape = (d["nl"] - 1, d["nod2D"])
        h = np.zeros(shape, order="F") if dttf_h is None else dttf_h
        v = np.zeros(shape, order="F") if dttf_v is None else dttf_v
        for a in (h, v):
            if a.shape != shape or a.dtype != np.float64 or not a.flags.f_contiguous:
                raise ValueError(f"tendency arrays must be F-contiguous float64 {shape}")
        _lib.check(self.lib.tracer_do_oce_adv_tra(tracer_id, float(dt), h, v, d["nl"], d["nod2D"]),
                   "tracer_do_oce_adv_tra", self.lib)
        return h, v

    def diagnostics(self, tracer_id=1):
        out = np.zeros(7)
        _lib.check(self.lib.tracer_get_diagnostics(tracer_id, out), "tracer_get_diagnostics", self.lib)
        keys = ("min", "max", "sum", "active_min", "active_max", "mass", "nan_count")
        return dict(zip(keys, out.tolist()))

    def work_array(self, name, count, tracer_id=1):
        out = np.zeros(count)
        _lib.check(self.lib.tracer_get_work_array(tracer_id, name.encode(), out, count), f"work array {name}",
                   self.lib)
        return out

    def time_steps(self, nsteps, dt):
        ms = ctypes.c_double()
        _lib.check(self.lib.tracer_time_steps(nsteps, dt, ctypes.byref(ms)), "tracer_time_steps", self.lib)
        return ms.value

    def time_kernels(self, dt, reps=3):
        names = ctypes.create_string_buffer(4096)
        ms = np.zeros(32)
        n = self.lib.tracer_time_kernels(dt, reps, names, 4096, ms, 32)
        if n < 0:
            raise RuntimeError(f"tracer_time_kernels failed: {_lib.last_error(self.lib)}")
        return dict(zip(names.value.decode().split(";"), ms[:n].tolist()))

    def precision_error(self, ref_precision, nsteps, dt):
        """Advance all tracers nsteps in lockstep with a second device state at
        ref_precision; returns (nsteps, ntracers, 4): max|a-b|, ||a-b||_2 and
        their relative forms against the reference state b."""
        ntr = self.num_tracers
        out = np.zeros(4 * nsteps * ntr)
        _lib.check(self.lib.tracer_precision_error(int(ref_precision), int(nsteps), float(dt), out),
                   "tracer_precision_error", self.lib)
        return out.reshape(nsteps, ntr, 4)

    def autotune(self, dt, reps=10):
        """Pick the fastest launch schedule on the current state (tracers are restored)."""
        table = np.zeros(3 * 32)
        n = self.lib.tracer_autotune(float(dt), int(reps), table, 32)
        if n < 0:
            raise RuntimeError(f"tracer_autotune failed: {_lib.last_error(self.lib)}")
        rows = table[:3 * n].reshape(n, 3)
        return [dict(block_waves=int(w), batch_tracers=int(b), ms_per_step=float(ms)) for w, b, ms in rows]

    def hbm_copy_rate(self, nbytes=4 << 30, reps=10):
        """Practical HBM ceiling: best (read + write) GB/s of the library's
        16-byte streaming copy over a few launch shapes, and that shape."""
        gbs, v = ctypes.c_double(), (ctypes.c_int * 3)()
        _lib.check(self.lib.tracer_hbm_copy_rate(int(nbytes), int(reps), ctypes.byref(gbs), v),
                   "tracer_hbm_copy_rate", self.lib)
        return dict(gbs=gbs.value, bytes_each_way=int(nbytes) // 16 * 16, workgroups=v[0], words_in_flight=v[1],
                    nontemporal=bool(v[2]))

    # ---- learned limiter
    def mlp_param_count(self):
        return int(self.lib.tracer_mlp_param_count())

    def mlp_params(self):
        p = np.zeros(self.mlp_param_count(), dtype=np.float32)
        _lib.check(self.lib.tracer_mlp_get_params(p, p.size), "tracer_mlp_get_params", self.lib)
        return p

    def set_mlp_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        _lib.check(self.lib.tracer_mlp_set_params(p, p.size), "tracer_mlp_set_params", self.lib)

    def learn_collect(self, nsteps, dt, max_rows):
        rows = ctypes.c_longlong()
        _lib.check(self.lib.tracer_learn_collect(int(nsteps), float(dt), int(max_rows), ctypes.byref(rows)),
                   "tracer_learn_collect", self.lib)
        return rows.value

    def learn_clear(self):
        self.lib.tracer_learn_clear()

    def learn_train(self, epochs, batch=65536, lr=1e-3, seed=0):
        out = np.zeros(int(epochs))
        _lib.check(self.lib.tracer_learn_train(int(epochs), int(batch), float(lr), int(seed), out),
                   "tracer_learn_train", self.lib)
        return out

    def mlp_forward(self, X):
        X = np.ascontiguousarray(X, dtype=np.float32)
        out = np.zeros(X.shape[0], dtype=np.float32)
        _lib.check(self.lib.tracer_mlp_forward(X, X.shape[0], out), "tracer_mlp_forward", self.lib)
        return out

    def mlp_gradient(self, X, y, scale=1.0):
        X = np.ascontiguousarray(X, dtype=np.float32)
        y = np.ascontiguousarray(y, dtype=np.float32)
        g = np.zeros(self.mlp_param_count(), dtype=np.float32)
        loss = ctypes.c_double()
        _lib.check(self.lib.tracer_mlp_gradient(X, y, X.shape[0], float(scale), g, ctypes.byref(loss)),
                   "tracer_mlp_gradient", self.lib)
        return g, loss.value

    def learn_fallback(self, tracer_id=1):
        f = ctypes.c_double()
        _lib.check(self.lib.tracer_learn_fallback(tracer_id, ctypes.byref(f)), "tracer_learn_fallback", self.lib)
        return f.value

    def cleanup(self):
        return self.lib.tracer_cleanup()
