"""The reference's pi mesh (FESOM mesh files, spherical geometry, coasts,
levels 5..46 of 48) on the GPU against the oracle, through the C ABI.

Bar (FP64): relative max-norm <= 1e-12 after each of 3 steps, mass conserved
to 1e-13, no new extrema with FCT beyond 1e-12 of the values' magnitude. FP32: mass to 1e-13, bounds to FP32
rounding.
"""
import numpy as np
import pytest

from conftest import mesh_dict

pytestmark = pytest.mark.gpu

SCHEMES = [("UPW1", "UPW1", "NONE"), ("UPW1", "UPW1", "FCT"), ("MUSCL", "QR4C", "FCT"), ("MFCT", "PPM", "FCT"),
           ("MUSCL", "CDIFF", "FCT"), ("MUSCL", "QR4C", "NONE")]


def _fields(m, L, ntr, seed):
    X = m["coord_nod2D"].reshape(-1, 2)
    lon, lat = X[:, 0], X[:, 1]
    rng = np.random.default_rng(seed)
    act = np.arange(L)[:, None] < (m["nlevels_nod2D"] - 1)[None, :]
    k = np.arange(L)[:, None]
    out = []
    for t in range(ntr):
        if t == 0:
            f = 15 + 5 * np.sin(lon)[None, :] * np.cos(2 * lat)[None, :] - 0.2 * k
        elif t == 1:
            f = 34 + np.exp(-((lon - rng.uniform(-3, 3)) ** 2 + (lat - rng.uniform(-1, 1)) ** 2) / 0.3)[None, :] + 0 * k
        elif t == 2:
            f = (np.sin(3 * lon) * np.cos(3 * lat) > 0).astype(float)[None, :] + 0 * k
        else:
            f = rng.uniform(0, 1, (L, len(lon)))
        out.append(np.asfortranarray(np.where(act, f, 0.0)))
    return out, act


def _setup(api, pi_mesh_dir, schemes, ntr=4, seed=2, prec=64):
    d = api.load_mesh_from_files(pi_mesh_dir)
    api.set_option("precision", prec)
    api.allocate_tracers(ntr)
    dt = api.make_velocity(0.1, 0.0, 0.2, 0.5, 0.3)
    m = mesh_dict(api)
    trs, act = _fields(m, d["nl"] - 1, ntr, seed)
    for i, t in enumerate(trs):
        api.set_values(i + 1, t)
        api.set_schemes(i + 1, *schemes)
    return d, m, dt, trs, act


@pytest.mark.parametrize("schemes", SCHEMES, ids=lambda s: "-".join(s))
def test_pi_mesh_fp64_matches_oracle(api, pi_mesh_dir, schemes):
    import oracle as orc
    print(f"[gpu] pi mesh parity {schemes}", flush=True)
    d, m, dt, trs, act = _setup(api, pi_mesh_dir, schemes)
    o = orc.Oracle(m)
    uv, w, we = api.mesh_array("uv"), api.mesh_array("w"), api.mesh_array("w_e")
    L, N = d["nl"] - 1, d["nod2D"]
    hn = m["hnode_new"].reshape(N, L).T
    av = m["areasvol"].reshape(N, d["nl"]).T[:L]
    ref = [np.array(t, order="F", copy=True) for t in trs]
    mass0 = [(t * hn * av)[act].sum() for t in ref]
    rng0 = [(t[act].min(), t[act].max()) for t in ref]
    for step in range(3):
        api.advect_steps(1, dt)
        for i in range(len(ref)):
            o.step(dt, uv, w, we, ref[i], hor=schemes[0], ver=schemes[1], lim=schemes[2])
            g = api.get_tracer_values(i + 1)
            err = np.max(np.abs(g - ref[i])) / np.max(np.abs(ref[i]))
            assert err <= 1e-12, f"tracer {i + 1} step {step + 1}: rel err {err:.3e}"
            mass = (g * hn * av)[act].sum()
            assert abs(mass - mass0[i]) <= 1e-13 * abs(mass0[i])
            if schemes[2] == "FCT":
                # W comes from discrete continuity on the sphere: a constant field
                # moves by ~3e-14 of its value per step, so the bound is relative
                # to the magnitude (S lies in 34..35), not to the range
                lo, hi = rng0[i]
                tol = 1e-12 * max(abs(lo), abs(hi))
                assert g[act].min() >= lo - tol and g[act].max() <= hi + tol


@pytest.mark.parametrize("prec", [32, 16])
def test_pi_mesh_reduced_precision(api, pi_mesh_dir, prec):
    print(f"[gpu] pi mesh precision {prec}", flush=True)
    d, m, dt, trs, act = _setup(api, pi_mesh_dir, ("MUSCL", "QR4C", "FCT"), prec=prec)
    L, N = d["nl"] - 1, d["nod2D"]
    hn = m["hnode_new"].reshape(N, L).T
    av = m["areasvol"].reshape(N, d["nl"]).T[:L]
    mass0 = [(t * hn * av)[act].sum() for t in trs]
    api.advect_steps(10, dt)
    for i, t in enumerate(trs):
        g = api.get_tracer_values(i + 1)
        assert abs((g * hn * av)[act].sum() - mass0[i]) <= 1e-13 * abs(mass0[i])
        lo, hi = t[act].min(), t[act].max()
        tol = 2.0 ** -23 * 10 * max(abs(lo), abs(hi))
        assert g[act].min() >= lo - tol and g[act].max() <= hi + tol
    api.set_option("precision", 64)


def test_pi_mesh_scalar_frames_follow_the_device_state(api, pi_mesh_dir, tmp_path):
    print("[gpu] pi mesh scalar output frames", flush=True)
    d, m, dt, trs, act = _setup(api, pi_mesh_dir, ("MUSCL", "QR4C", "FCT"), ntr=2)
    api.output_open(tmp_path, 2)
    api.output_step()
    want = [[t.T.copy() for t in trs]]
    for _ in range(2):
        api.advect_steps(1, dt)
        api.output_step()
        want.append([api.get_tracer_values(i + 1).T.copy() for i in range(2)])
    api.output_close()
    L = d["nl"] - 1
    for i, name in enumerate(["temperature", "salinity"]):
        frames = np.fromfile(tmp_path / f"{name}.bin", dtype=np.float64).reshape((3, d["nod2D"], L))
        for f in range(3):
            assert np.array_equal(frames[f], want[f][i])
