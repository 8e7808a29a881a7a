"""Batched form: several meshes advanced by one graph launch per step.

The captured multi-mesh graph must give bitwise the same tracers as eager
per-mesh launches, conserve every tracer's mass (FP64 accumulation) and keep
FCT's bounds on every mesh.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MESHES = [dict(nx=40, ny=32, nl=12, Lx=4.0e5, Ly=3.2e5, seed=1, ntracers=6, precision=32),
          dict(nx=24, ny=24, nl=20, Lx=2.4e5, Ly=2.4e5, seed=2, ntracers=5, precision=64),
          dict(nx=56, ny=40, nl=16, Lx=5.6e5, Ly=4.0e5, seed=3, ntracers=3, precision=32)]


def _build():
    from fesom2_tracer_dwarf_amd.batch import TracerBatch
    b = TracerBatch()
    for kw in MESHES:
        b.add_periodic_mesh(kw["nx"], kw["ny"], kw["nl"], kw["Lx"], kw["Ly"], seed=kw["seed"],
                            ntracers=kw["ntracers"], precision=kw["precision"])
    return b


def test_batch_graph_matches_eager_and_conserves():
    print("[gpu] batch: 3 meshes, graph vs eager", flush=True)
    b = _build()
    d0 = [[b.diagnostics(i, t + 1) for t in range(kw["ntracers"])] for i, kw in enumerate(MESHES)]
    b.run(4, use_graph=False)
    eager = [[b.values(i, t + 1) for t in range(kw["ntracers"])] for i, kw in enumerate(MESHES)]
    b.close()
    b = _build()
    b.run(4, use_graph=True)
    for i, kw in enumerate(MESHES):
        for t in range(kw["ntracers"]):
            g = b.values(i, t + 1)
            assert np.array_equal(g, eager[i][t]), f"mesh {i} tracer {t + 1}: graph differs from eager"
            d1 = b.diagnostics(i, t + 1)
            m0, m1 = d0[i][t]["mass"], d1["mass"]
            assert abs(m1 - m0) <= 1e-13 * abs(m0), f"mesh {i} tracer {t + 1}: mass drift"
            lo, hi = d0[i][t]["active_min"], d0[i][t]["active_max"]
            tol = (2.0 ** -23 if kw["precision"] == 32 else 1e-12) * 4 * max(abs(lo), abs(hi))
            assert d1["active_min"] >= lo - tol and d1["active_max"] <= hi + tol
    ms = b.time(3)
    assert ms > 0.0
    b.close()


def test_batch_schedules_give_identical_tracers():
    """The batch autotuner (each mesh picks its own waves per workgroup and
    tracers per launch) and a fixed schedule change only the launch shape:
    the tracers after a few steps are bitwise those of the default schedule."""
    print("[gpu] batch: default vs autotuned vs (2 waves, 1 tracer) schedules", flush=True)
    results = []
    for mode in ("default", "tuned", "fixed"):
        b = _build()
        if mode == "tuned":
            sched = b.autotune(1)
            assert len(sched) == len(MESHES)
            assert all(w in (2, 4, 8) and nt in (1, 2, 4) for w, nt in sched)
        elif mode == "fixed":
            b.set_schedule(2, 1)
        b.run(3, use_graph=True)
        results.append([[b.values(i, t + 1) for t in range(kw["ntracers"])] for i, kw in enumerate(MESHES)])
        b.close()
    for other in results[1:]:
        for i, kw in enumerate(MESHES):
            for t in range(kw["ntracers"]):
                assert np.array_equal(other[i][t], results[0][i][t]), f"mesh {i} tracer {t + 1}"
