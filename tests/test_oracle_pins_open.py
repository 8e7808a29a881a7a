"""Pins of the oracle's PPM limiter, fill_up_dn_grad, CDIFF and MUSCL's c_lo
masking, each against something other than its own formula.

* PPM (lib/oce_adv_tra_ver.F90:448-644): on uniform layers the reference's
  interface formula reduces to Colella & Woodward (1984) eq. 1.6 with the
  monotonised slope of eq. 1.8, and its limiter is their eq. 1.10. The
  restatement below is written from that paper's uniform-zone form; it builds
  each cell's parabola from its three conditions (edge values aL, aR and the
  cell mean) by a linear solve and integrates it over the swept part of the
  cell by Gauss-Legendre quadrature, so neither the limiter's closed forms
  (aL = 3T - 2aR ...) nor the Courant-integrated flux formula is retyped.
  Random columns reach all four limiter branches (counted).
* fill_up_dn_grad (lib/oce_muscl_adv.F90:361-530): a vectorised restatement
  of its definition -- the up/down-wind triangle's gradient on the levels all
  elements around both nodes share, the elem_area-weighted mean over the
  node's wet elements elsewhere -- on the irregular wheel mesh (unequal
  spherical triangles, random bottoms), where an unweighted mean or a wrong
  level split gives different numbers (checked).
* CDIFF (lib/oce_adv_tra_ver.F90:648-712): a linear profile on uniform layers
  has its interface value exactly at the mean of the two layers around it.
* MUSCL c_lo (lib/oce_adv_tra_hor.F90:364-365): below nboundary_lay of a node
  the reconstruction at that node is its own value, so the node's gradients
  must not change the flux there, must change it above, and where both ends
  are masked the flux has the closed form
  -1/2 (1 - phi) [(v + |v|) T1 + (v - |v|) T2] - phi v (T1 + T2) / 2,
  with the transport v read off two UPW1 fluxes of unit node values.
"""
import ctypes

import numpy as np
import pytest

import oracle as orc
from conftest import mesh_dict, write_wheel_mesh

P = lambda x: x.ctypes.data_as(ctypes.POINTER(ctypes.c_double))  # noqa: E731


# ---------------------------------------------------------------------------
# PPM
# ---------------------------------------------------------------------------
def _cw84_slope(T, j):
    """Monotonised slope of cell j on a uniform grid (Colella & Woodward 1984, eq. 1.8)."""
    d = 0.5 * (T[j + 1] - T[j - 1])
    if (T[j + 1] - T[j]) * (T[j] - T[j - 1]) > 0.0:
        return np.sign(d) * min(abs(d), 2.0 * abs(T[j + 1] - T[j]), 2.0 * abs(T[j] - T[j - 1]))
    return 0.0


def _parabola_mean(aL, aR, mean, p, q):
    """Mean over [p, q] of the parabola on [0, 1] with P(0) = aL, P(1) = aR and
    cell mean `mean` (two-point Gauss-Legendre, exact for quadratics)."""
    c = np.linalg.solve(np.array([[1.0, 0.0, 0.0], [1.0, 1.0, 1.0], [1.0, 0.5, 1.0 / 3.0]]),
                        np.array([aL, aR, mean]))
    m, r = 0.5 * (p + q), 0.5 * (q - p) / np.sqrt(3.0)
    Pf = lambda x: c[0] + c[1] * x + c[2] * x * x  # noqa: E731
    return 0.5 * (Pf(m - r) + Pf(m + r))


def ppm_column(T, W, area, h, dt, branches):
    """Interface fluxes (L+1) of one column of L uniform layers."""
    L = T.size
    a = np.empty(L + 1)
    a[0] = T[0]
    a[1] = 0.5 * (T[0] + T[1])
    for i in range(2, L - 1):                                   # CW84 eq. 1.6, uniform zones
        a[i] = 0.5 * (T[i - 1] + T[i]) - (_cw84_slope(T, i) - _cw84_slope(T, i - 1)) / 6.0
    a[L - 1] = T[L - 1] if W[L - 1] >= 0.0 else T[L - 2]        # the upwind value above the bottom layer
    a[L] = T[L - 1]
    F = np.zeros(L + 1)
    for j in range(L):
        if W[j] <= 0.0 and W[j + 1] >= 0.0:                     # nothing leaves this layer
            continue
        aL, aR, t = a[j], a[j + 1], T[j]
        s = (aR - aL) * (t - 0.5 * (aL + aR))
        if (aR - t) * (t - aL) <= 0.0:                          # CW84 eq. 1.10: local extremum
            aL = aR = t
            branches["flat"] += 1
        elif s > (aR - aL) ** 2 / 6.0:
            aL = 3.0 * t - 2.0 * aR
            branches["left"] += 1
        elif s < -(aR - aL) ** 2 / 6.0:
            aR = 3.0 * t - 2.0 * aL
            branches["right"] += 1
        else:
            branches["smooth"] += 1
        if W[j] > 0.0:                                           # up through the top: the top x of the cell
            x = min(W[j] * dt / h[j], 1.0)
            F[j] = -W[j] * _parabola_mean(aL, aR, t, 0.0, x) * area[j]
        if W[j + 1] < 0.0:                                       # down through the bottom
            x = min(-W[j + 1] * dt / h[j], 1.0)
            F[j + 1] = -W[j + 1] * _parabola_mean(aL, aR, t, 1.0 - x, 1.0) * area[j + 1]
    F[0] = -a[0] * W[0] * area[0]
    F[L] = 0.0
    return F


def test_ppm_limiter_against_colella_woodward(api):
    nl = 14
    d = api.generate_mesh(8, 6, nl, periodic=True, Lx=8e4, Ly=6e4, dz=10.0, stretch=0.0, lev_lo=nl, lev_hi=nl,
                          seed=0)
    m = mesh_dict(api)
    o = orc.Oracle(m)
    N, L = d["nod2D"], nl - 1
    rng = np.random.default_rng(11)
    T = np.asfortranarray(rng.uniform(0.0, 1.0, (L, N)))
    # a few columns built to sit in a branch: monotone ramps with a step, smooth profiles
    T[:, 0] = np.linspace(0.0, 1.0, L)
    T[:, 1] = np.where(np.arange(L) < 6, 0.0, 1.0)
    T[:, 2] = np.array([0, 0, 0.1, 1, 1, 1, 0.9, 0.2, 0.15, 0.1, 0.1, 0, 0])[:L]
    T[:, 3] = np.sin(np.arange(L) * 0.4)
    dt = 600.0
    W = np.asfortranarray(rng.normal(0.0, 8e-3, (nl, N)))       # Courant numbers ~0.5, some clipped at 1
    W[np.abs(W) < 1e-6] = 1e-6
    flux = np.zeros((nl, N), order="F")
    o.lib.oracle_adv_tra_vert_ppm(ctypes.byref(o.mesh), dt, P(W), P(T), P(flux), 1)
    area = m["area"].reshape(N, nl).T
    h = m["hnode"].reshape(N, L).T
    branches = dict(flat=0, left=0, right=0, smooth=0)
    ref = np.stack([ppm_column(T[:, n], W[:, n], area[:, n], h[:, n], dt, branches) for n in range(N)], axis=1)
    scale = np.abs(ref).max()
    assert np.abs(flux - ref).max() <= 1e-12 * scale
    assert min(branches.values()) >= 20, branches
    assert np.any(np.abs(W) * dt / h[0, 0] > 1.0)               # the Courant clip is reached too


def test_ppm_extremum_flux_is_upwind_donor_value(api):
    """At a local extremum the parabola is flat, so the flux is -W T area of
    the donor layer exactly."""
    nl = 10
    d = api.generate_mesh(6, 5, nl, periodic=True, Lx=6e4, Ly=5e4, dz=10.0, stretch=0.0, lev_lo=nl, lev_hi=nl,
                          seed=0)
    m = mesh_dict(api)
    o = orc.Oracle(m)
    N, L = d["nod2D"], nl - 1
    T = np.asfortranarray(np.tile(np.array([1.0, 2.0, 1.0, 3.0, 0.5, 2.5, 0.2, 1.7, 0.9])[:L, None], (1, N)))
    W = np.asfortranarray(np.tile(np.array([1e-3, 2e-3, -3e-3, 1.5e-3, -1e-3, 2e-3, -2e-3, 1e-3, -1e-3, 0.0])
                                  [:, None], (1, N)))
    flux = np.zeros((nl, N), order="F")
    o.lib.oracle_adv_tra_vert_ppm(ctypes.byref(o.mesh), 300.0, P(W), P(T), P(flux), 1)
    area = m["area"].reshape(N, nl).T
    # layers 1..6 (0-based) alternate maxima and minima; interface k's donor is layer k (W > 0) or k-1 (W < 0)
    for k in range(2, 7):
        donor = k if W[k, 0] > 0 else k - 1
        assert np.allclose(flux[k], -W[k] * T[donor] * area[k], rtol=1e-14)


# ---------------------------------------------------------------------------
# fill_up_dn_grad
# ---------------------------------------------------------------------------
@pytest.fixture(scope="module")
def wheel_mesh(api, tmp_path_factory):
    dd = tmp_path_factory.mktemp("wheel_pins")
    write_wheel_mesh(dd, spokes=(12, 9), nx=24, ny=18, nl=12, seed=8)
    api.load_mesh_from_files(str(dd))
    return mesh_dict(api)


def _node_means(m, trxy):
    """elem_area-weighted mean of the elemental gradients over each node's elements that hold the level."""
    N, T, L = m["nod2D"], m["elem2D"], m["nl"] - 1
    nie = m["nod_in_elem2D"].reshape(N, m["maxAdj"])
    cnt = m["nod_in_elem2D_num"]
    ea = m["elem_area"]
    wet = np.arange(L)[None, :] < (m["nlevels"] - 1)[:, None]          # (T, L)
    num = np.zeros((N, L, 2))
    den = np.zeros((N, L))
    for k in range(m["maxAdj"]):
        has = k < cnt
        el = np.where(has, nie[:, k], 0)
        wk = (wet[el] & has[:, None]) * ea[el][:, None]                    # (N, L)
        num += wk[:, :, None] * trxy[el]
        den += wk
    with np.errstate(invalid="ignore", divide="ignore"):
        return num / den[:, :, None]


def _fill_up_dn_grad(m, trxy, mean):
    N, E, L = m["nod2D"], m["edge2D"], m["nl"] - 1
    ed = m["edges"].reshape(E, 2)
    ud = m["edge_up_dn_tri"].reshape(E, 2)
    lev = np.arange(L)[None, :]
    out = np.full((E, L, 4), np.nan)
    for side in (0, 1):
        n = ed[:, side]
        wet = lev < (m["nlevels_nod2D"][n] - 1)[:, None]
        g = mean[n]                                                          # (E, L, 2)
        for comp in (0, 1):
            out[:, :, side + 2 * comp] = np.where(wet, g[:, :, comp], np.nan)
    both = (ud[:, 0] >= 0) & (ud[:, 1] >= 0)
    nzmax = np.minimum(m["nlevels_nod2D_min"][ed[:, 0]], m["nlevels_nod2D_min"][ed[:, 1]])
    shared = both[:, None] & (lev < (nzmax - 1)[:, None])                   # levels 1 .. nzmax-1
    up, dn = np.maximum(ud[:, 0], 0), np.maximum(ud[:, 1], 0)
    for c, (tri, comp) in enumerate(((up, 0), (dn, 0), (up, 1), (dn, 1))):
        out[:, :, c] = np.where(shared, trxy[tri][:, :, comp], out[:, :, c])
    return out, shared, both


def test_fill_up_dn_grad_restatement_on_irregular_mesh(wheel_mesh):
    m = wheel_mesh
    o = orc.Oracle(m)
    N, T, E, L = m["nod2D"], m["elem2D"], m["edge2D"], m["nl"] - 1
    X = m["coord_nod2D"].reshape(N, 2)
    rng = np.random.default_rng(4)
    ttf = np.asfortranarray(np.sin(3.0 * X[:, 0])[None, :] * np.cos(5.0 * X[:, 1])[None, :]
                            + rng.uniform(0.0, 0.3, (L, N)))
    trxy = np.zeros(2 * L * T)
    o.lib.oracle_tracer_gradient_elements(ctypes.byref(o.mesh), P(ttf), P(trxy))
    grad = np.zeros(4 * L * E)
    o.lib.oracle_fill_up_dn_grad(ctypes.byref(o.mesh), P(trxy), P(grad))
    trxy = trxy.reshape(T, L, 2)
    grad = grad.reshape(E, L, 4)
    ref, shared, both = _fill_up_dn_grad(m, trxy, _node_means(m, trxy))
    defined = ~np.isnan(ref)
    scale = np.abs(ref[defined]).max()
    assert np.abs(grad[defined] - ref[defined]).max() <= 1e-12 * scale
    # both branches reached: triangle gradients on shared levels, nodal means below them and on coast edges
    below = both[:, None] & ~shared & defined[:, :, 0]
    assert shared.sum() > 1000 and below.sum() > 100 and (~both).sum() > 10
    # the test sees the weighting: an unweighted mean would differ
    flat = dict(m)
    flat["elem_area"] = np.ones(T)
    ref_flat, _, _ = _fill_up_dn_grad(m, trxy, _node_means(flat, trxy))
    mean_levels = defined & ~shared[:, :, None]
    assert np.abs(ref_flat[mean_levels] - ref[mean_levels]).max() > 1e-6 * scale


# ---------------------------------------------------------------------------
# CDIFF
# ---------------------------------------------------------------------------
def test_cdiff_exact_interface_values_of_linear_profiles(api):
    nl = 12
    d = api.generate_mesh(6, 5, nl, periodic=True, Lx=6e4, Ly=5e4, dz=10.0, stretch=0.0, lev_lo=3, lev_hi=nl,
                          seed=2)
    m = mesh_dict(api)
    o = orc.Oracle(m)
    N = d["nod2D"]
    Z = m["Z_3d_n"].reshape(N, nl - 1).T
    zb = m["zbar_3d_n"].reshape(N, nl).T
    area = m["area"].reshape(N, nl).T
    a, b = 3.0, 0.02
    T = np.asfortranarray(a + b * Z)
    W = np.asfortranarray(np.random.default_rng(5).normal(0.0, 1e-4, (nl, N)))
    flux = np.zeros((nl, N), order="F")
    o.lib.oracle_adv_tra_ver_cdiff(ctypes.byref(o.mesh), P(W), P(T), P(flux), 1)
    for n in range(N):
        L = m["nlevels_nod2D"][n] - 1
        k = np.arange(1, L)                                     # interior interfaces
        exact = -(a + b * zb[k, n]) * W[k, n] * area[k, n]
        assert np.allclose(flux[k, n], exact, rtol=1e-12, atol=1e-15 * np.abs(exact).max())
        assert flux[0, n] == -W[0, n] * T[0, n] * area[0, n]
        assert flux[L, n] == 0.0


# ---------------------------------------------------------------------------
# MUSCL c_lo masking at nboundary_lay
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("phi", [0.0, 1.0, 0.35])
def test_muscl_clo_masking(api, phi):
    nl = 9
    d = api.generate_mesh(12, 10, nl, periodic=False, Lx=11e3, Ly=9e3, dz=10.0, lev_lo=3, lev_hi=nl, seed=2)
    m = mesh_dict(api)
    o = orc.Oracle(m)
    N, E, T, L = d["nod2D"], d["edge2D"], d["elem2D"], nl - 1
    rng = np.random.default_rng(7)
    ttf = np.asfortranarray(rng.uniform(1.0, 2.0, (L, N)))
    uv = np.asfortranarray(rng.normal(0.0, 0.1, (2, L, T)))
    trxy = np.zeros(2 * L * T)
    o.lib.oracle_tracer_gradient_elements(ctypes.byref(o.mesh), P(ttf), P(trxy))
    grad = np.zeros(4 * L * E)
    o.lib.oracle_fill_up_dn_grad(ctypes.byref(o.mesh), P(trxy), P(grad))

    def muscl(g):
        f = np.zeros(L * E)
        o.lib.oracle_adv_tra_hor_muscl(ctypes.byref(o.mesh), P(np.ravel(uv, order="F")), P(ttf), phi, P(f),
                                       P(np.ascontiguousarray(g)), 1)
        return f.reshape(E, L)

    def upw1(t):
        f = np.zeros(L * E)
        o.lib.oracle_adv_tra_hor_upw1(ctypes.byref(o.mesh), P(np.ravel(uv, order="F")), P(np.asfortranarray(t)),
                                      P(f), 1)
        return f.reshape(E, L)

    F0 = muscl(grad)
    ed = m["edges"].reshape(E, 2)
    nbl = m["nboundary_lay"]
    lev = np.arange(1, L + 1)[None, :]
    et = m["edge_tri"].reshape(E, 2)
    nlev_e = np.maximum(m["nlevels"][et[:, 0]] - 1, np.where(et[:, 1] >= 0, m["nlevels"][np.maximum(et[:, 1], 0)] - 1, 0))
    wet = lev <= nlev_e[:, None]
    masked = [(lev > nbl[ed[:, s]][:, None]) & wet for s in (0, 1)]
    # the transport, read off the UPW1 flux of T = 1 (-1/2 [(v + |v|) + (v - |v|)] = -v exactly)
    v = -upw1(np.ones((L, N)))
    for side in (0, 1):
        g = grad.reshape(E, L, 4).copy()
        g[:, :, side] += rng.normal(0.0, 1e-3, (E, L))          # this end's x and y gradients
        g[:, :, side + 2] += rng.normal(0.0, 1e-3, (E, L))
        F1 = muscl(g.ravel())
        assert np.array_equal(F1[masked[side]], F0[masked[side]])
        # ... and they do matter above nboundary_lay (with phi = 0 only on the upwind end)
        upwind = (v > 0) if side == 0 else (v < 0)
        live = wet & ~masked[side] & (upwind if phi == 0.0 else True)
        assert live.sum() > 100 and np.mean(F1[live] != F0[live]) > 0.9
    # both ends masked: the closed form
    both = masked[0] & masked[1]
    assert both.sum() > 100
    T1 = ttf.T[ed[:, 0]]                                          # (E, L)
    T2 = ttf.T[ed[:, 1]]
    Fu = upw1(ttf)                                                # -1/2 [(v+|v|) T1 + (v-|v|) T2]
    closed = -0.5 * (1.0 - phi) * ((v + np.abs(v)) * T1 + (v - np.abs(v)) * T2) - phi * v * 0.5 * (T1 + T2)
    scale = np.abs(F0[wet]).max()
    assert np.abs(F0[both] - closed[both]).max() <= 1e-13 * scale
    if phi == 0.0:
        assert np.abs(F0[both] - Fu[both]).max() <= 1e-15 * scale
