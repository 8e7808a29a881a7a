"""16-wide node records on the GPU: a Delaunay mesh with hubs of 12 and 16
edges (FESOM mesh files, spherical geometry) against the oracle, through the
C ABI. Seven tracers with one scheme run as batches of 4, 2 and 1, so every
batch width of every node kernel is exercised at record width 16.

Bars: FP64 relative max-norm <= 1e-12 after each of 3 steps, mass to 1e-13,
no new extrema with FCT beyond 1e-12 of the values' magnitude; FP32 <= 1e-5
and FP16 gradients <= 1e-3 against the oracle after 3 steps; the learned
limiter conserves mass to 1e-13 and creates no new extrema (FP64 1e-12, FP32
2^-23 per step, relative to the magnitude). The same mesh handed over as
arrays by a caller steps bit for bit as the file-loaded one.
"""
import numpy as np
import pytest

from conftest import mesh_dict, write_wheel_mesh

pytestmark = pytest.mark.gpu

NTR = 7


@pytest.fixture(scope="module")
def wheel_dir(tmp_path_factory):
    d = tmp_path_factory.mktemp("wheel")
    deg = write_wheel_mesh(d, spokes=(12, 16))
    assert deg.max() == 16
    return d


def _fields(m, L, seed):
    X = m["coord_nod2D"].reshape(-1, 2)
    lon, lat = X[:, 0], X[:, 1]
    rng = np.random.default_rng(seed)
    act = np.arange(L)[:, None] < (m["nlevels_nod2D"] - 1)[None, :]
    k = np.arange(L)[:, None]
    c = X.mean(axis=0)
    out = []
    for t in range(NTR):
        if t % 4 == 0:
            f = 15 + 5 * np.sin(40 * lon)[None, :] * np.cos(60 * lat)[None, :] - 0.2 * k
        elif t % 4 == 1:
            f = 34 + np.exp(-((lon - c[0] - rng.uniform(-0.02, 0.02)) ** 2 + (lat - c[1]) ** 2) / 4e-4)[None, :] + 0 * k
        elif t % 4 == 2:
            f = (np.sin(80 * lon) * np.cos(80 * lat) > 0).astype(float)[None, :] + 0 * k
        else:
            f = rng.uniform(0, 1, (L, len(lon)))
        out.append(np.asfortranarray(np.where(act, f, 0.0)))
    return out, act


def _setup(api, wheel_dir, schemes, seed=3):
    d = api.load_mesh_from_files(wheel_dir)
    api.set_option("precision", 64)
    api.allocate_tracers(NTR)
    dt = api.make_velocity(0.1, 0.0, 0.2, 0.5, 0.3)
    m = mesh_dict(api)
    trs, act = _fields(m, d["nl"] - 1, seed)
    for i, t in enumerate(trs):
        api.set_values(i + 1, t)
        api.set_schemes(i + 1, *schemes)
    L, N = d["nl"] - 1, d["nod2D"]
    hn = m["hnode_new"].reshape(N, L).T
    av = m["areasvol"].reshape(N, d["nl"]).T[:L]
    return d, m, dt, trs, act, hn * av


SCHEMES = [("MUSCL", "QR4C", "FCT"), ("MFCT", "PPM", "FCT"), ("UPW1", "CDIFF", "FCT"), ("MUSCL", "QR4C", "NONE"),
           ("MFCT", "UPW1", "NONE")]


@pytest.mark.parametrize("schemes", SCHEMES, ids=lambda s: "-".join(s))
def test_wide_records_fp64_match_oracle(api, wheel_dir, schemes):
    import oracle as orc
    print(f"[gpu] 16-wide node records, FP64 {schemes}", flush=True)
    d, m, dt, trs, act, vol = _setup(api, wheel_dir, schemes)
    o = orc.Oracle(m)
    uv, w, we = api.mesh_array("uv"), api.mesh_array("w"), api.mesh_array("w_e")
    ref = [np.array(t, order="F", copy=True) for t in trs]
    mass0 = [(t * vol)[act].sum() for t in ref]
    rng0 = [(t[act].min(), t[act].max()) for t in ref]
    for step in range(3):
        api.advect_steps(1, dt)
        for i in range(NTR):
            o.step(dt, uv, w, we, ref[i], hor=schemes[0], ver=schemes[1], lim=schemes[2])
            g = api.get_tracer_values(i + 1)
            err = np.max(np.abs(g - ref[i])) / np.max(np.abs(ref[i]))
            assert err <= 1e-12, f"tracer {i + 1} step {step + 1}: rel err {err:.3e}"
            assert abs((g * vol)[act].sum() - mass0[i]) <= 1e-13 * abs(mass0[i])
            if schemes[2] == "FCT":
                lo, hi = rng0[i]
                tol = 1e-12 * max(abs(lo), abs(hi))
                assert g[act].min() >= lo - tol and g[act].max() <= hi + tol
    assert np.abs(api.get_tracer_values(1) - trs[0]).max() > 1e-6


@pytest.mark.parametrize("prec,bar", [(32, 1e-5), (16, 1e-3)], ids=["fp32", "fp16grad"])
def test_wide_records_reduced_precision(api, wheel_dir, prec, bar):
    import oracle as orc
    schemes = ("MUSCL", "QR4C", "FCT")
    print(f"[gpu] 16-wide node records, precision {prec}", flush=True)
    d, m, dt, trs, act, vol = _setup(api, wheel_dir, schemes)
    o = orc.Oracle(m)
    uv, w, we = api.mesh_array("uv"), api.mesh_array("w"), api.mesh_array("w_e")
    api.set_option("precision", prec)
    try:
        ref = [np.array(t, order="F", copy=True) for t in trs]
        mass0 = [(t * vol)[act].sum() for t in ref]
        for step in range(3):
            api.advect_steps(1, dt)
            for i in range(NTR):
                o.step(dt, uv, w, we, ref[i], hor=schemes[0], ver=schemes[1], lim=schemes[2])
        for i in range(NTR):
            g = api.get_tracer_values(i + 1)
            err = np.max(np.abs(g - ref[i])) / np.max(np.abs(ref[i]))
            print(f"[gpu]   tracer {i + 1}: rel err {err:.2e}", flush=True)
            assert err <= bar, f"tracer {i + 1}: rel err {err:.3e}"
            assert abs((g * vol)[act].sum() - mass0[i]) <= 1e-13 * abs(mass0[i])
    finally:
        api.set_option("precision", 64)


@pytest.mark.parametrize("prec", [64, 32])
def test_wide_records_learned_limiter(api, wheel_dir, prec):
    print(f"[gpu] 16-wide node records, learned limiter, precision {prec}", flush=True)
    d, m, dt, trs, act, vol = _setup(api, wheel_dir, ("MUSCL", "QR4C", "LEARNED"))
    api.set_option("precision", prec)
    try:
        mass0 = [(t * vol)[act].sum() for t in trs]
        api.advect_steps(3, dt)
        for i in range(NTR):
            g = api.get_tracer_values(i + 1)
            assert np.all(np.isfinite(g))
            assert abs((g * vol)[act].sum() - mass0[i]) <= 1e-13 * abs(mass0[i]), f"tracer {i + 1}: mass drift"
            lo, hi = trs[i][act].min(), trs[i][act].max()
            tol = (1e-12 if prec == 64 else 2.0 ** -23 * 3) * max(abs(lo), abs(hi))
            assert g[act].min() >= lo - tol and g[act].max() <= hi + tol, f"tracer {i + 1}: new extrema"
    finally:
        api.set_option("precision", 64)


def test_wide_records_through_caller_supplied_mesh(api, wheel_dir):
    """The same mesh handed over as arrays (tracer_mesh_begin / tracer_set_mesh_array / tracer_mesh_end, required
    arrays plus coordinates and up/down-wind triangles) steps bit for bit as the file-loaded one."""
    from test_mesh_arrays import REQUIRED
    print("[gpu] 16-wide node records, caller-supplied mesh", flush=True)
    schemes = ("MUSCL", "QR4C", "FCT")
    d, m, dt, trs, act, vol = _setup(api, wheel_dir, schemes)
    given = {k: api.mesh_array(k) for k in REQUIRED + ["edge_up_dn_tri", "coord_nod2D", "areasvol", "hnode_new"]}
    uv, w, we = api.mesh_array("uv"), api.mesh_array("w"), api.mesh_array("w_e")
    api.advect_steps(2, dt)
    ref = [api.get_tracer_values(i + 1) for i in range(NTR)]
    got = api.set_mesh(d["nl"], d["nod2D"], d["elem2D"], d["edge2D"], d["maxAdj"], given)
    assert got == d
    api.allocate_tracers(NTR)
    L = d["nl"] - 1
    api.set_velocity_elem(uv.reshape(2, L, d["elem2D"], order="F"), w.reshape(d["nl"], d["nod2D"], order="F"),
                          we.reshape(d["nl"], d["nod2D"], order="F"))
    for i, t in enumerate(trs):
        api.set_values(i + 1, t)
        api.set_schemes(i + 1, *schemes)
    api.advect_steps(2, dt)
    for i in range(NTR):
        g = api.get_tracer_values(i + 1)
        assert np.array_equal(g.view(np.uint64), ref[i].view(np.uint64)), f"tracer {i + 1}"
