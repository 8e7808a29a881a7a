"""Reduced-precision forms against FP64: FP32 fluxes/gradients/limiter with the
node tracer and divergence accumulation kept in FP64, and the same with the
gradients stored as FP16.

Bars (written here, measured through the C ABI):
  FP32      relative max-norm <= 1e-5 against the FP64 GPU path after every
            step of 20, and <= 1e-5 against the CPU FP64 oracle after 3 steps;
  FP16 grad relative max-norm <= 1e-3 against FP64 after every step of 20;
  both      volume-weighted mass conserved to 1e-13 relative (the fluxes enter
            two nodes with opposite sign in FP64), and with FCT no value
            outside the initial range by more than 2^-23 * max|T| per step.
            The LO solution, the bounds and the accumulation are FP64, but the
            velocity and transports are stored in FP32, so the discrete
            divergence of the flow is zero only to FP32 rounding; a tracer
            with a large mean (S ~ 34, span 1) then drifts outside its range
            by about that much per step, and FP64 stays exact.
Tracers: T (smooth, 0..20), S (34 + blob), a 0/1 step function, a seeded
Gaussian blob and white noise in [0,1] (the limiter's hard cases),
MUSCL + QR4C + FCT.
"""
import numpy as np
import pytest

from conftest import mesh_dict
from test_gpu_parity import rel_err, setup_case

pytestmark = pytest.mark.gpu

MESH = dict(nx=64, ny=48, nl=24, periodic=True, Lx=3.2e5, Ly=2.4e5, dz=10.0, stretch=0.5, lev_lo=3, lev_hi=24,
            seed=11)
SCHEMES = ("MUSCL", "QR4C", "FCT")
BARS = {32: 1e-5, 16: 1e-3}
MASS_TOL = 1e-13
NTR = 5


def _mass_and_range(api, m, d, ntr):
    L, N = d["nl"] - 1, d["nod2D"]
    hn = m["hnode_new"].reshape(N, L).T
    av = m["areasvol"].reshape(N, d["nl"]).T[:L]
    act = np.arange(L)[:, None] < (m["nlevels_nod2D"] - 1)[None, :]
    out = []
    for i in range(ntr):
        g = api.get_tracer_values(i + 1)
        out.append(((g * hn * av)[act].sum(), g[act].min(), g[act].max()))
    return out


@pytest.mark.parametrize("prec", [32, 16])
def test_reduced_precision_against_fp64(api, prec):
    print(f"[gpu] precision {prec} vs 64, 20 steps", flush=True)
    d, dt, trs = setup_case(api, MESH, SCHEMES, ntr=NTR, seed=2)
    m = mesh_dict(api)
    before = _mass_and_range(api, m, d, NTR)
    api.set_option("precision", prec)
    err = api.precision_error(64, 20, dt)
    assert err.shape == (20, NTR, 4)
    worst = err[:, :, 2].max()
    print(f"[gpu] FP{prec}: worst relative max-norm {worst:.3e}; per tracer after 20 steps "
          f"{err[-1, :, 2].tolist()}", flush=True)
    assert np.all(np.isfinite(err))
    assert worst <= BARS[prec], f"relative max-norm {worst:.3e} above {BARS[prec]:.0e}"
    after = _mass_and_range(api, m, d, NTR)
    for i, ((m0, lo, hi), (m1, lo1, hi1)) in enumerate(zip(before, after)):
        assert abs(m1 - m0) <= MASS_TOL * abs(m0), f"tracer {i + 1}: mass drift {abs(m1 - m0) / abs(m0):.3e}"
        tol = 2.0 ** -23 * max(abs(lo), abs(hi)) * err.shape[0]
        assert lo1 >= lo - tol and hi1 <= hi + tol, f"tracer {i + 1}: new extrema {lo - lo1:.3e} {hi1 - hi:.3e}"
    api.set_option("precision", 64)


def test_fp32_against_oracle(api):
    import oracle as orc
    print("[gpu] FP32 vs oracle, 3 steps", flush=True)
    d, dt, trs = setup_case(api, MESH, SCHEMES, ntr=4, seed=2)
    m = mesh_dict(api)
    o = orc.Oracle(m)
    uv, w, we = api.mesh_array("uv"), api.mesh_array("w"), api.mesh_array("w_e")
    api.set_option("precision", 32)
    ref = [np.array(t, order="F", copy=True) for t in trs]
    for step in range(3):
        api.advect_steps(1, dt)
        for i in range(4):
            o.step(dt, uv, w, we, ref[i], hor=SCHEMES[0], ver=SCHEMES[1], lim=SCHEMES[2])
            err = rel_err(api.get_tracer_values(i + 1), ref[i])
            assert err <= BARS[32], f"tracer {i + 1} step {step + 1}: rel err {err:.3e}"
    api.set_option("precision", 64)


def test_precision_error_of_fp64_against_itself_is_zero(api):
    d, dt, trs = setup_case(api, MESH, SCHEMES, ntr=2, seed=2)
    err = api.precision_error(64, 3, dt)
    assert np.all(err == 0.0)


def test_fp32_limiter_factors_against_fp64(api):
    """Zalesak's R+/R- (work arrays fct_plus / fct_minus) after one step from
    the same state: in [0, 1] in both precisions (min(1, Q/P) with Q and P of
    one sign), and the FP32 factors -- which the device keeps interleaved per
    cell pair -- equal the FP64 ones to rounding except where the limiter
    sits on a switching point: a swapped or shifted read-out would differ on
    the limited cells, which the second bar looks at."""
    print("[gpu] FP32 R+/R- against FP64", flush=True)
    d, dt, trs = setup_case(api, MESH, SCHEMES, ntr=2)
    m = mesh_dict(api)
    L, N = d["nl"] - 1, d["nod2D"]
    act = np.arange(L)[None, :] < (m["nlevels_nod2D"] - 1)[:, None]   # (N, L)
    out = {}
    for prec in (64, 32):
        api.set_option("precision", prec)
        for i, t in enumerate(trs):
            api.set_values(i + 1, t)
        api.advect_steps(1, dt)
        out[prec] = [api.work_array(k, L * N, tracer_id=1).reshape(N, L)[act] for k in ("fct_plus", "fct_minus")]
    for name, a64, a32 in zip(("R+", "R-"), out[64], out[32]):
        for a in (a64, a32):
            assert a.min() >= 0.0 and a.max() <= 1.0, name
        diff = np.abs(a32 - a64)
        lim = a64 < 0.5
        far = np.mean(diff > 1e-2)
        med = np.median(diff[lim]) if lim.any() else 0.0
        print(f"[gpu]   {name}: {lim.sum()} limited layers, median |d| there {med:.2e}, share |d| > 1e-2 {far:.2e}",
              flush=True)
        assert lim.sum() > 0, name
        assert far <= 1e-2 and med <= 1e-4, name


def test_fp32_edge_fluxes_against_fp64(api):
    """The antidiffusive edge fluxes (work array adv_flux_hor, HO - LO) of
    both tracers of a two-tracer launch after one step from the same state.
    The device keeps FP32 fluxes of two work slots interleaved per cell pair,
    so tracer 2 reads the second half. Tracer 2 is tracer 1 doubled: HO and LO
    are linear in the tracer and doubling is exact in floating point, so its
    fluxes are exactly twice tracer 1's in both precisions; tracer 1's FP32
    fluxes equal the FP64 ones to a small relative error where there is flux."""
    print("[gpu] FP32 HO-LO edge fluxes: both work slots, against FP64", flush=True)
    d, dt, trs = setup_case(api, MESH, SCHEMES, ntr=2)
    trs[1] = 2.0 * trs[0]
    m = mesh_dict(api)
    L, E = d["nl"] - 1, d["edge2D"]
    el = m["edge_tri"].reshape(E, 2)
    nlv = m["nlevels"]
    nle = np.maximum(nlv[el[:, 0]], np.where(el[:, 1] >= 0, nlv[np.maximum(el[:, 1], 0)], 0)) - 1
    act = np.arange(L)[None, :] < nle[:, None]                       # (E, L) wet edge levels
    out = {}
    for prec in (64, 32):
        api.set_option("precision", prec)
        for i, t in enumerate(trs):
            api.set_values(i + 1, t)
        api.advect_steps(1, dt)
        out[prec] = [api.work_array("adv_flux_hor", L * E, tracer_id=i + 1).reshape(E, L)[act] for i in range(2)]
        assert np.array_equal(out[prec][1], 2.0 * out[prec][0]), f"FP{prec}: slot 2 is not slot 1 doubled"
    a64, a32 = out[64][0], out[32][0]
    big = np.abs(a64) > 1e-3 * np.abs(a64).max()
    rel = np.abs(a32 - a64)[big] / np.abs(a64)[big]
    print(f"[gpu]   {big.sum()} edge levels with flux, median relative FP32 difference {np.median(rel):.2e}",
          flush=True)
    assert big.sum() > 0 and np.median(rel) <= 1e-4
