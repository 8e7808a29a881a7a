"""Pins of the CPU oracle (and of the mesh setup it shares with the GPU path)
against closed forms, invariants and independent restatements.

Each check is chosen so that a plausible transcription mistake in the oracle
(a sign, a swapped index, a dropped term, a wrong level range) fails it.
"""
import numpy as np
import pytest

import oracle as orc
from conftest import mesh_dict

R_EARTH = 6367500.0


def closed_mesh(api, nx=12, ny=10, nl=6, Lx=11e3, Ly=9e3, flat=True, seed=0, sfc=False):
    lo = nl if flat else 3
    d = api.generate_mesh(nx, ny, nl, periodic=False, Lx=Lx, Ly=Ly, dz=10.0, lev_lo=lo, lev_hi=nl, seed=seed,
                          sfc=sfc)
    return d, mesh_dict(api)


def periodic_mesh(api, nx=16, ny=12, nl=10, Lx=1.6e5, Ly=1.2e5, stretch=0.5, flat=False, seed=3):
    lo = nl if flat else 3
    d = api.generate_mesh(nx, ny, nl, periodic=True, Lx=Lx, Ly=Ly, dz=10.0, stretch=stretch, lev_lo=lo, lev_hi=nl,
                          seed=seed, sfc=True)
    return d, mesh_dict(api)


def uniform_velocity(d, U, V):
    L, T = d["nl"] - 1, d["elem2D"]
    uv = np.zeros((2, L, T), order="F")
    uv[0], uv[1] = U, V
    return uv


def interior_nodes(X, Lx, Ly, margin):
    return (X[:, 0] > margin) & (X[:, 0] < Lx - margin) & (X[:, 1] > margin) & (X[:, 1] < Ly - margin)


# ---------------------------------------------------------------------------
# geometry and topology of the setup (mesh_areas / mesh_auxiliary_arrays)
# ---------------------------------------------------------------------------
def test_periodic_geometry_closed_forms(api):
    nx, ny, Lx, Ly = 16, 12, 1.6e5, 1.2e5
    d, m = periodic_mesh(api, nx, ny, nl=6, Lx=Lx, Ly=Ly, stretch=0.0, flat=True)
    dx, dy = Lx / nx, Ly / ny
    assert np.allclose(m["elem_area"], 0.5 * dx * dy, rtol=1e-12)
    area = m["area"].reshape(d["nod2D"], d["nl"])
    assert np.allclose(area[:, :-1], dx * dy, rtol=1e-12)     # 6 triangles / 3 per node, every layer
    assert np.all(area[:, -1] == 0.0)                         # no element has the bottom interface as a layer
    assert np.all(m["nod_in_elem2D_num"] == 6)                # complete adjacency at the periodic wrap
    assert d["edge2D"] == 3 * d["nod2D"] and d["elem2D"] == 2 * d["nod2D"]
    # centroid-minus-midpoint components are multiples of dx/6, dy/6 with |.| <= dx/3, dy/3
    ecd = m["edge_cross_dxdy"].reshape(-1, 4)
    for c, h in ((0, dx), (1, dy), (2, dx), (3, dy)):
        q = ecd[:, c] / (h / 6.0)
        assert np.allclose(q, np.round(q), atol=1e-9) and np.all(np.abs(np.round(q)) <= 2)


def test_orientation_and_clockwise(api):
    d, m = closed_mesh(api, 9, 7)
    X = api.mesh_array("coord").reshape(-1, 2)
    en = m["elem2D_nodes"].reshape(-1, 3)
    b, c = X[en[:, 1]] - X[en[:, 0]], X[en[:, 2]] - X[en[:, 0]]
    assert np.all(b[:, 0] * c[:, 1] - b[:, 1] * c[:, 0] < 0)  # clockwise, as test_tri enforces
    ed, et = m["edges"].reshape(-1, 2), m["edge_tri"].reshape(-1, 2)
    cen = X[en].mean(axis=1)
    e_vec = X[ed[:, 1]] - X[ed[:, 0]]
    r1 = cen[et[:, 0]] - X[ed[:, 0]]
    assert np.all(e_vec[:, 0] * r1[:, 1] - e_vec[:, 1] * r1[:, 0] > 0)  # edge_tri(1) on the left
    inner = et[:, 1] >= 0
    r2 = cen[et[inner, 1]] - X[ed[inner, 0]]
    assert np.all(e_vec[inner, 0] * r2[:, 1] - e_vec[inner, 1] * r2[:, 0] < 0)
    # boundary edges: exactly the rim
    nb = (~inner).sum()
    assert nb == 2 * (9 - 1) + 2 * (7 - 1)


def test_gradient_sca_exact_on_linear_fields(api):
    d, m = periodic_mesh(api, 11, 9, nl=5, Lx=1.1e5, Ly=0.9e5, flat=True)
    X = api.mesh_array("coord").reshape(-1, 2)
    en = m["elem2D_nodes"].reshape(-1, 3)
    g = m["gradient_sca"].reshape(-1, 6)
    Lx, Ly = 1.1e5, 0.9e5
    rel = X[en] - X[en[:, :1]]
    rel[..., 0] -= Lx * np.round(rel[..., 0] / Lx)
    rel[..., 1] -= Ly * np.round(rel[..., 1] / Ly)
    for a, b in ((1.0, 0.0), (0.0, 1.0), (0.37, -1.3)):
        f = 4.0 + a * rel[..., 0] + b * rel[..., 1]
        assert np.allclose((g[:, :3] * f).sum(1), a, atol=1e-12 * (abs(a) + abs(b) + 1))
        assert np.allclose((g[:, 3:] * f).sum(1), b, atol=1e-12 * (abs(a) + abs(b) + 1))


def test_levels_follow_fesom_invariant(api):
    d, m = periodic_mesh(api, 14, 10, nl=12)
    en = m["elem2D_nodes"].reshape(-1, 3)
    nn = m["nlevels_nod2D"]
    nie = m["nod_in_elem2D"].reshape(d["nod2D"], d["maxAdj"])
    for n in range(d["nod2D"]):
        els = nie[n, :m["nod_in_elem2D_num"][n]]
        assert nn[n] == m["nlevels"][els].max()
        assert m["nlevels_nod2D_min"][n] == m["nlevels"][els].min()
    assert np.all(m["nlevels"] <= nn[en].min(axis=1))
    assert nn.min() >= 3 and nn.max() <= d["nl"]


def test_up_downwind_and_nboundary_restatements_agree(api):
    for per in (True, False):
        if per:
            d, m = periodic_mesh(api, 13, 9, nl=8, Lx=1.3e5, Ly=0.9e5)
            cx, cy = 1.3e5 / R_EARTH, 0.9e5 / R_EARTH
        else:
            d, m = closed_mesh(api, 10, 8, nl=8, flat=False, seed=4)
            cx = cy = 0.0
        o = orc.Oracle(m)
        assert np.array_equal(o.up_downwind(cx, cy), m["edge_up_dn_tri"])
        assert np.array_equal(o.nboundary_lay(), m["nboundary_lay"])
    # closed mesh: rim nodes have nboundary_lay 0, interior nodes >= 1
    X = api.mesh_array("coord").reshape(-1, 2)
    rim = (X[:, 0] == X[:, 0].min()) | (X[:, 0] == X[:, 0].max()) | (X[:, 1] == X[:, 1].min()) | \
          (X[:, 1] == X[:, 1].max())
    assert np.all(m["nboundary_lay"][rim] == 0) and np.all(m["nboundary_lay"][~rim] >= 1)


def test_upwind_triangle_contains_the_upwind_direction(api):
    d, m = closed_mesh(api, 9, 8, nl=4)
    X = api.mesh_array("coord").reshape(-1, 2)
    en = m["elem2D_nodes"].reshape(-1, 3)
    ed = m["edges"].reshape(-1, 2)
    ud = m["edge_up_dn_tri"].reshape(-1, 2)
    for e in range(d["edge2D"]):
        for side, node, direction in ((0, ed[e, 0], X[ed[e, 0]] - X[ed[e, 1]]), (1, ed[e, 1], X[ed[e, 1]] - X[ed[e, 0]])):
            t = ud[e, side]
            if t < 0:
                continue
            others = [q for q in en[t] if q != node]
            b, c = X[others[0]] - X[node], X[others[1]] - X[node]
            # direction lies in the closed cone spanned by b and c
            cb = b[0] * direction[1] - b[1] * direction[0]
            cc = c[0] * direction[1] - c[1] * direction[0]
            assert cb * cc <= 1e-9 * np.linalg.norm(b) * np.linalg.norm(direction) ** 2


# ---------------------------------------------------------------------------
# horizontal schemes: exact transport of linear fields by a uniform current
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("hor,ph", [("UPW1", 1.0), ("MUSCL", 1.0), ("MUSCL", 0.0), ("MFCT", 0.5)])
def test_linear_field_uniform_flow_exact(api, hor, ph):
    """With FESOM's orientation every scheme gives dT/dt = -u.grad T exactly
    at interior nodes (an edge_tri(1)-on-the-right mesh gives +1/3 instead)."""
    d, m = closed_mesh(api)
    o = orc.Oracle(m)
    X = api.mesh_array("coord").reshape(-1, 2)
    L, N = d["nl"] - 1, d["nod2D"]
    a, b, U, V = 0.3e-3, -0.2e-3, 0.11, -0.07
    T0 = np.asfortranarray(np.repeat((5 + a * X[:, 0] + b * X[:, 1])[None, :], L, axis=0))
    uv = uniform_velocity(d, U, V)
    w = np.zeros((d["nl"], N), order="F")
    v = T0.copy(order="F")
    dt = 10.0
    o.step(dt, uv, w, w, v, hor=hor, ver="UPW1", lim="NONE", ph=ph)
    inner = interior_nodes(X, 11e3, 9e3, 1.5e3)
    err = np.abs((v - T0)[:, inner] / dt + (U * a + V * b)).max()
    assert err <= 1e-12 * abs(U * a + V * b)


def test_orientation_mistake_is_detected(api):
    """Plant the reference analytic generator's orientation error (edge_tri(1)
    on the right for horizontal/vertical edges) in the oracle's inputs: the
    exact-transport pin must fail."""
    d, m = closed_mesh(api)
    X = api.mesh_array("coord").reshape(-1, 2)
    ed = m["edges"].reshape(-1, 2).copy()
    horiz_or_vert = (X[ed[:, 0], 0] == X[ed[:, 1], 0]) | (X[ed[:, 0], 1] == X[ed[:, 1], 1])
    inner = m["edge_tri"].reshape(-1, 2)[:, 1] >= 0
    flip = horiz_or_vert & inner
    ed[flip] = ed[flip][:, ::-1]
    ecd = m["edge_cross_dxdy"].reshape(-1, 4)
    bad = dict(m)
    bad["edges"] = ed.ravel()
    o = orc.Oracle(bad)
    L, N = d["nl"] - 1, d["nod2D"]
    T0 = np.asfortranarray(np.repeat((5 + 0.3e-3 * X[:, 0])[None, :], L, axis=0))
    v = T0.copy(order="F")
    w = np.zeros((d["nl"], N), order="F")
    o.step(10.0, uniform_velocity(d, 0.1, 0.0), w, w, v, hor="UPW1", ver="UPW1", lim="NONE")
    inner_n = interior_nodes(X, 11e3, 9e3, 1.5e3)
    err = np.abs((v - T0)[:, inner_n] / 10.0 + 0.1 * 0.3e-3).max()
    assert err > 1e-2 * 0.1 * 0.3e-3
    assert ecd.shape[0] == d["edge2D"]


# ---------------------------------------------------------------------------
# vertical schemes on linear-in-z profiles
# ---------------------------------------------------------------------------
def _column_setup(api, stretch):
    d, m = periodic_mesh(api, 6, 5, nl=12, Lx=6e4, Ly=5e4, stretch=stretch, flat=True)
    o = orc.Oracle(m)
    N, nl = d["nod2D"], d["nl"]
    Z = m["Z_3d_n"].reshape(N, nl - 1).T
    zb = m["zbar_3d_n"].reshape(N, nl).T
    area = m["area"].reshape(N, nl).T
    return d, m, o, Z, zb, area


@pytest.mark.parametrize("num_ord", [1.0, 0.0, 0.4])
@pytest.mark.parametrize("stretch", [0.0, 0.8])
def test_qr4c_exact_for_linear_profiles(api, num_ord, stretch):
    import ctypes
    d, m, o, Z, zb, area = _column_setup(api, stretch)
    N, nl = d["nod2D"], d["nl"]
    a, b, Wc = 3.0, 0.02, 1.3e-4
    T = np.asfortranarray(a + b * Z)
    W = np.full((nl, N), Wc, order="F")
    flux = np.zeros((nl, N), order="F")
    p = lambda x: x.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    o.lib.oracle_adv_tra_ver_qr4c(ctypes.byref(o.mesh), p(W), p(T), num_ord, p(flux), 1)
    ks = np.arange(2, nl - 2)  # interfaces 3..nlev-2 (0-based 2..nl-3)
    exact = -(a + b * zb[ks]) * Wc * area[ks]
    assert np.allclose(flux[ks], exact, rtol=1e-12, atol=1e-14 * np.abs(exact).max())
    assert np.all(flux[nl - 1] == 0.0)                                  # zero bottom flux
    assert np.allclose(flux[0], -T[0] * Wc * area[0], rtol=1e-14)       # surface upwind


def test_ppm_exact_for_linear_profiles_on_uniform_layers(api):
    import ctypes
    d, m, o, Z, zb, area = _column_setup(api, 0.0)
    N, nl = d["nod2D"], d["nl"]
    a, b, Wc, dt = 3.0, 0.02, 2.0e-3, 600.0
    T = np.asfortranarray(a + b * Z)
    W = np.full((nl, N), Wc, order="F")
    flux = np.zeros((nl, N), order="F")
    p = lambda x: x.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    o.lib.oracle_adv_tra_vert_ppm(ctypes.byref(o.mesh), dt, p(W), p(T), p(flux), 1)
    h = m["hnode"].reshape(N, nl - 1).T
    for j in range(2, nl - 3):          # interfaces whose both bounding values come from the PPM formula
        aL, aR = a + b * zb[j], a + b * zb[j + 1]
        x = min(Wc * dt / h[j, 0], 1.0)
        exact = (-aL - 0.5 * x * (aR - aL)) * area[j] * Wc
        assert np.allclose(flux[j], exact, rtol=1e-12)


def test_qr4c_two_layer_column_keeps_input_flux(api):
    """Reference quirk: on a two-layer column interface nzmin+1 is written twice
    (lib/oce_adv_tra_ver.F90:405-411), so it returns the input flux."""
    import ctypes
    d = api.generate_mesh(6, 5, 3, periodic=True, Lx=6e4, Ly=5e4, dz=10.0, lev_lo=3, lev_hi=3, seed=0)
    m = mesh_dict(api)
    o = orc.Oracle(m)
    N = d["nod2D"]
    T = np.asfortranarray(np.random.default_rng(0).uniform(size=(2, N)))
    W = np.asfortranarray(np.random.default_rng(1).normal(size=(3, N)) * 1e-4)
    fin = np.asfortranarray(np.random.default_rng(2).normal(size=(3, N)))
    flux = fin.copy(order="F")
    p = lambda x: x.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    o.lib.oracle_adv_tra_ver_qr4c(ctypes.byref(o.mesh), p(W), p(T), 1.0, p(flux), 0)
    cen = -0.5 * (T[0] + T[1]) * W[1] * m["area"].reshape(N, 3)[:, 1]
    assert np.array_equal(flux[1], cen - (cen - fin[1]))
    assert np.all(np.abs(flux[1] - fin[1]) <= 4e-16 * np.abs(cen) + 1e-15)


# ---------------------------------------------------------------------------
# the full step: invariants
# ---------------------------------------------------------------------------
SCHEMES = [("UPW1", "UPW1", "NONE"), ("UPW1", "UPW1", "FCT"), ("MUSCL", "QR4C", "FCT"), ("MFCT", "PPM", "FCT"),
           ("MUSCL", "CDIFF", "FCT"), ("MUSCL", "QR4C", "NONE"), ("MFCT", "CDIFF", "NONE")]


def _setup_flow(api, periodic=True):
    if periodic:
        d, m = periodic_mesh(api)
    else:
        d, m = closed_mesh(api, 12, 10, nl=9, flat=False, seed=2)
    dt = api.make_velocity(0.1, 0.05, 0.2, 0.5, 0.3)
    uv, w = api.mesh_array("uv"), api.mesh_array("w")
    return d, m, dt, uv, w


@pytest.mark.parametrize("schemes", SCHEMES, ids=lambda s: "-".join(s))
@pytest.mark.parametrize("periodic", [True, False])
def test_constant_field_preserved(api, schemes, periodic):
    """T == c stays c: checks the discrete continuity of w against the edge
    fluxes, the vertical fluxes and the FCT sequence together."""
    d, m, dt, uv, w = _setup_flow(api, periodic)
    o = orc.Oracle(m)
    v = np.asfortranarray(np.full((d["nl"] - 1, d["nod2D"]), 7.25))
    for _ in range(3):
        o.step(dt, uv, w, w, v, hor=schemes[0], ver=schemes[1], lim=schemes[2])
    L = d["nl"] - 1
    act = np.arange(L)[:, None] < (m["nlevels_nod2D"] - 1)[None, :]
    assert np.abs(v[act] - 7.25).max() <= 1e-12


def test_no_flow_no_change(api):
    d, m = periodic_mesh(api)
    o = orc.Oracle(m)
    rng = np.random.default_rng(3)
    v = np.asfortranarray(rng.uniform(size=(d["nl"] - 1, d["nod2D"])))
    v0 = v.copy()
    uv = np.zeros(2 * (d["nl"] - 1) * d["elem2D"])
    w = np.zeros(d["nl"] * d["nod2D"])
    o.step(100.0, uv, w, w, v, hor="MUSCL", ver="QR4C", lim="FCT")
    assert np.array_equal(v, v0)


@pytest.mark.parametrize("schemes", SCHEMES, ids=lambda s: "-".join(s))
def test_mass_conserved_and_fct_bounded(api, schemes):
    from fesom2_tracer_dwarf_amd import configs
    d, m, dt, uv, w = _setup_flow(api)
    o = orc.Oracle(m)
    L, N = d["nl"] - 1, d["nod2D"]
    hn = m["hnode_new"].reshape(N, L).T
    av = m["areasvol"].reshape(N, d["nl"]).T[:L]
    act = np.arange(L)[:, None] < (m["nlevels_nod2D"] - 1)[None, :]
    for t in configs.initial_tracers(api.mesh_array("coord"), d["nl"], 1.6e5, 1.2e5, 4, seed=5):
        v = np.array(t, order="F", copy=True)
        m0, lo, hi = (v * hn * av)[act].sum(), v[act].min(), v[act].max()
        for _ in range(6):
            o.step(dt, uv, w, w, v, hor=schemes[0], ver=schemes[1], lim=schemes[2])
        assert abs((v * hn * av)[act].sum() - m0) <= 1e-13 * abs(m0)
        if schemes[2] == "FCT":
            assert v[act].min() >= lo - 1e-12 * (hi - lo) and v[act].max() <= hi + 1e-12 * (hi - lo)


def test_fct_with_equal_ho_and_lo_is_the_lo_solution(api):
    d, m, dt, uv, w = _setup_flow(api)
    o = orc.Oracle(m)
    rng = np.random.default_rng(7)
    v1 = np.asfortranarray(rng.uniform(size=(d["nl"] - 1, d["nod2D"])))
    v2 = v1.copy(order="F")
    o.step(dt, uv, w, w, v1, hor="UPW1", ver="UPW1", lim="FCT")
    o.step(dt, uv, w, w, v2, hor="UPW1", ver="UPW1", lim="NONE")
    assert np.abs(v1 - v2).max() <= 1e-13 * np.abs(v2).max()


def test_step_function_stays_in_unit_interval(api):
    from fesom2_tracer_dwarf_amd import configs
    d, m, dt, uv, w = _setup_flow(api)
    o = orc.Oracle(m)
    rng = np.random.default_rng(1)
    v = configs.tracer_step(api.mesh_array("coord"), d["nl"] - 1, 1.6e5, 1.2e5, rng)
    act = np.arange(d["nl"] - 1)[:, None] < (m["nlevels_nod2D"] - 1)[None, :]
    for _ in range(10):
        o.step(dt, uv, w, w, v, hor="MUSCL", ver="QR4C", lim="FCT")
    assert v[act].min() >= -1e-14 and v[act].max() <= 1 + 1e-14
    assert 0.0 < v[act].mean() < 1.0


# ---------------------------------------------------------------------------
# an independent vectorised Zalesak limiter against the oracle's loops
# ---------------------------------------------------------------------------
def numpy_zalesak(m, d, dt, ttf, lo, adf_h, adf_v):
    """Zalesak (1979) FCT with FESOM's cluster bounds, written from the
    algorithm (not the loops): bounds over the elements around each node and
    the layers above and below, P+/P- sums, R+/R-, donor/receiver limiting."""
    L, N, nl = d["nl"] - 1, d["nod2D"], d["nl"]
    nl_n = m["nlevels_nod2D"] - 1
    nl_e = m["nlevels"] - 1
    lev = np.arange(L)
    act = lev[:, None] < nl_n[None, :]
    big = 1e3
    a1max, a1min = np.maximum(lo, ttf), np.minimum(lo, ttf)
    en = m["elem2D_nodes"].reshape(-1, 3)
    wet_e = lev[:, None] < nl_e[None, :]
    emax = np.where(wet_e, a1max[:, en].max(axis=2), -big)
    emin = np.where(wet_e, a1min[:, en].min(axis=2), big)
    nie = m["nod_in_elem2D"].reshape(N, d["maxAdj"])
    cnt = m["nod_in_elem2D_num"]
    tvmax = np.full((L, N), -np.inf)
    tvmin = np.full((L, N), np.inf)
    for j in range(d["maxAdj"]):
        ok = j < cnt
        el = np.where(ok, nie[:, j], 0)
        tvmax = np.where(ok[None, :], np.maximum(tvmax, emax[:, el]), tvmax)
        tvmin = np.where(ok[None, :], np.minimum(tvmin, emin[:, el]), tvmin)
    up_max = np.vstack([tvmax[:1], tvmax[:-1]])
    dn_max = np.vstack([tvmax[1:], tvmax[-1:]])
    up_min = np.vstack([tvmin[:1], tvmin[:-1]])
    dn_min = np.vstack([tvmin[1:], tvmin[-1:]])
    edge_layer = (lev[:, None] == 0) | (lev[:, None] == nl_n[None, :] - 1)
    qmax = np.where(edge_layer, tvmax, np.maximum(np.maximum(up_max, tvmax), dn_max)) - lo
    qmin = np.where(edge_layer, tvmin, np.minimum(np.minimum(up_min, tvmin), dn_min)) - lo
    pp = np.maximum(0, adf_v[:L]) + np.maximum(0, -adf_v[1:])
    pm = np.minimum(0, adf_v[:L]) + np.minimum(0, -adf_v[1:])
    ed = m["edges"].reshape(-1, 2)
    et = m["edge_tri"].reshape(-1, 2)
    nle = np.maximum(nl_e[et[:, 0]], np.where(et[:, 1] >= 0, nl_e[np.maximum(et[:, 1], 0)], 0))
    on = lev[:, None] < nle[None, :]
    fh = np.where(on, adf_h, 0.0)
    for col, sgn in ((0, 1.0), (1, -1.0)):
        np.add.at(pp.T, ed[:, col], np.maximum(0, sgn * fh).T)
        np.add.at(pm.T, ed[:, col], np.minimum(0, sgn * fh).T)
    av = m["areasvol"].reshape(N, nl).T[:L]
    with np.errstate(divide="ignore", invalid="ignore"):  # dry levels have areasvol 0
        rp = np.minimum(1.0, qmax / (pp * dt / av + 1e-16))
        rm = np.minimum(1.0, qmin / (pm * dt / av - 1e-16))
    rp, rm = np.where(act, rp, 0), np.where(act, rm, 0)
    # vertical: interface k between layers k-1 (above) and k (below)
    fv = adf_v.copy()
    ae0 = np.where(fv[0] >= 0, np.minimum(1, rp[0]), np.minimum(1, rm[0]))
    fv[0] = ae0 * fv[0]
    for k in range(1, L):
        ok = k <= nl_n - 1
        ae = np.where(fv[k] >= 0, np.minimum(np.minimum(1, rm[k - 1]), rp[k]),
                      np.minimum(np.minimum(1, rp[k - 1]), rm[k]))
        fv[k] = np.where(ok, ae * fv[k], fv[k])
    n1, n2 = ed[:, 0], ed[:, 1]
    ae = np.where(adf_h >= 0, np.minimum(np.minimum(1, rp[:, n1]), rm[:, n2]),
                  np.minimum(np.minimum(1, rm[:, n1]), rp[:, n2]))
    fhl = np.where(on, ae * adf_h, adf_h)
    return fhl, fv


def test_fct_against_independent_numpy_zalesak(api):
    import ctypes
    d, m, dt, uv, w = _setup_flow(api)
    o = orc.Oracle(m)
    L, N, E, nl = d["nl"] - 1, d["nod2D"], d["edge2D"], d["nl"]
    rng = np.random.default_rng(11)
    act = np.arange(L)[:, None] < (m["nlevels_nod2D"] - 1)[None, :]
    ttf = np.asfortranarray(np.where(act, rng.uniform(0, 1, (L, N)), 0.0))
    lo = np.asfortranarray(np.where(act, ttf + rng.normal(0, 0.05, (L, N)), 0.0))
    nl_e = m["nlevels"] - 1
    et = m["edge_tri"].reshape(-1, 2)
    nle = np.maximum(nl_e[et[:, 0]], np.where(et[:, 1] >= 0, nl_e[np.maximum(et[:, 1], 0)], 0))
    on = np.arange(L)[:, None] < nle[None, :]
    adf_h = np.asfortranarray(np.where(on, rng.normal(0, 2e4, (L, E)), 0.0))
    iface = np.arange(nl)[:, None] < (m["nlevels_nod2D"] - 1)[None, :]
    iface[0] = True
    adf_v = np.asfortranarray(np.where(iface, rng.normal(0, 2e4, (nl, N)), 0.0))
    adf_v[0] = rng.normal(0, 2e4, N)
    fh_ref, fv_ref = numpy_zalesak(m, d, dt, ttf, lo, adf_h, adf_v)
    fh, fv = adf_h.copy(order="F"), adf_v.copy(order="F")
    p = lambda x: x.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    o.lib.oracle_tra_adv_fct(ctypes.byref(o.mesh), dt, p(ttf), p(lo), p(fh), p(fv), ctypes.byref(o.work))
    assert np.allclose(fh, fh_ref, rtol=1e-13, atol=1e-9)
    assert np.allclose(fv, fv_ref, rtol=1e-13, atol=1e-9)
    # the limiter actually bites somewhere
    assert np.abs(fh - adf_h).max() > 1e-3 * np.abs(adf_h).max()
