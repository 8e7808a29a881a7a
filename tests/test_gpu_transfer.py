"""Tracer hand-over between host arrays and the device (tracer_set_values,
tracer_get_values, tracer_set_values_ab) at a size that spans two of the
64 MB pinned staging chunks (420 x 400 nodes x 63 layers: 85 MB per tracer).

  * set then get returns every value bit for bit: wet layers through the
    device packing, the levels below each column's bottom through the
    device copy of the dense array; NaN payloads, -0.0 and subnormals kept;
  * after a step, the levels below the bottom are still exactly as set, the
    wet layers moved, and the all-values diagnostics equal numpy's min / max
    (exact) and sum (1e-13) over the returned array;
  * a get of the first nn < nod2D columns returns that prefix of the full get;
  * values uploaded with tracer_set_values_ab in Adams-Bashforth mode reach
    the step: a step from (values, valuesAB) equals the step from
    (values, values) only when the two are equal.
"""
import ctypes

import numpy as np
import pytest

from conftest import mesh_dict

pytestmark = pytest.mark.gpu

MESH = dict(periodic=True, Lx=4.2e5, Ly=4.0e5, dz=10.0, stretch=0.3, lev_lo=3, lev_hi=64, seed=9, sfc=True)


def _setup(api, ntr=2):
    d = api.generate_mesh(420, 400, 64, **MESH)
    api.set_option("precision", 64)
    api.allocate_tracers(ntr)
    dt = api.make_velocity(0.1, 0.05, 0.2, 0.5, 0.3)
    m = mesh_dict(api)
    L, N = d["nl"] - 1, d["nod2D"]
    act = np.arange(L)[:, None] < (m["nlevels_nod2D"] - 1)[None, :]
    assert L * N * 8 > 64 << 20
    return d, dt, act


def _values(act, seed):
    rng = np.random.default_rng(seed)
    t = np.asfortranarray(np.where(act, 10.0 + rng.standard_normal(act.shape), 1000.0 + rng.uniform(0, 1, act.shape)))
    dry = np.argwhere(~act)
    special = np.array([0x7FF8000000000123, 0x8000000000000000, 0x0000000000000001, 0x800FFFFFFFFFFFFF],
                       dtype=np.uint64).view(np.float64)  # NaN payload, -0.0, smallest subnormals
    for q, (k, n) in enumerate(dry[[5, 1000, 20000, -1]]):
        t[k, n] = special[q]
    return t


def _bits(a):
    return np.ascontiguousarray(a).view(np.uint64)


def test_round_trip_bit_exact_and_dry_levels_kept(api):
    print("[gpu] host hand-over, 420x400x64 (two staging chunks)", flush=True)
    d, dt, act = _setup(api)
    vals = [_values(act, s) for s in (1, 2)]
    for i, v in enumerate(vals):
        api.set_values(i + 1, v)
        api.set_schemes(i + 1, "MUSCL", "QR4C", "FCT")
    for i, v in enumerate(vals):
        assert np.array_equal(_bits(api.get_tracer_values(i + 1)), _bits(v)), f"tracer {i + 1}"
    api.advect_steps(2, dt)
    for i, v in enumerate(vals):
        g = api.get_tracer_values(i + 1)
        assert np.array_equal(_bits(g[~act]), _bits(v[~act])), f"tracer {i + 1}: levels below the bottom changed"
        assert np.all(np.isfinite(g[act])) and np.abs(g[act] - v[act]).max() > 1e-6
        dg = api.diagnostics(i + 1)
        ok = ~np.isnan(g)
        assert dg["nan_count"] == (~ok).sum() == 1
        assert dg["min"] == g[ok].min() and dg["max"] == g[ok].max()
        assert abs(dg["sum"] - g[ok].sum()) <= 1e-13 * np.abs(g[ok]).sum()
        assert dg["active_min"] == g[act].min() and dg["active_max"] == g[act].max()


def test_partial_get_is_a_prefix(api):
    print("[gpu] tracer_get_values_id with nn < nod2D", flush=True)
    d, dt, act = _setup(api, ntr=1)
    v = _values(act, 3)
    api.set_values(1, v)
    api.set_schemes(1, "UPW1", "UPW1", "NONE")
    api.advect_steps(1, dt)
    full = api.get_tracer_values(1)
    L, nn = d["nl"] - 1, 12345
    part = np.zeros((L, nn), order="F")
    assert api.lib.tracer_get_values_id(1, part, L, nn) == 0
    assert np.array_equal(_bits(part), _bits(full[:, :nn]))


def test_values_ab_upload_reaches_the_step(api):
    print("[gpu] tracer_set_values_ab in Adams-Bashforth mode", flush=True)
    d, dt, act = _setup(api, ntr=1)
    v = _values(act, 4)
    ab = np.asfortranarray(np.where(act, v + 0.25 * np.sin(np.arange(act.shape[1]))[None, :], v))
    api.set_schemes(1, "MUSCL", "QR4C", "FCT")
    res = {}
    try:
        api.set_option("refresh_valuesAB", 0)
        for tag, a in (("same", v), ("ab", ab), ("same_again", v)):
            api.set_values(1, v)
            api.set_values_ab(1, a)
            api.advect_steps(1, dt)
            res[tag] = api.get_tracer_values(1)
    finally:
        api.set_option("refresh_valuesAB", 1)
    api.set_values(1, v)
    api.advect_steps(1, dt)
    refresh = api.get_tracer_values(1)
    assert np.array_equal(_bits(res["same"]), _bits(res["same_again"]))
    assert np.array_equal(_bits(res["same"]), _bits(refresh))
    assert np.abs(res["ab"][act] - res["same"][act]).max() > 1e-6
