"""The drop-in boundary on the GPU, through the C ABI:

* tracer_do_oce_adv_tra returns what do_oce_adv_tra returns
  (lib/oce_adv_tra_driver.F90:23-294): the horizontal and vertical tendencies
  ADDED into the caller's arrays, the tracer unchanged -- against the
  oracle's do_oce_adv_tra, relative max-norm <= 1e-12 of each tendency;
* the accumulate_del_ttf option reproduces the reference C interface's
  tracer_advect_step, whose del_ttf is never zeroed
  (lib/tracer_c_interface.F90:176-212), step after step against the oracle
  run the same way (1e-12);
* the Fortran module fortran/oce_adv_tra_driver.F90, called by a FESOM-style
  host program with FESOM's derived types (tests/fortran), returns the same
  tendencies as the oracle, and adds them onto what the arrays held.
"""
import subprocess
from pathlib import Path

import numpy as np
import pytest

from conftest import mesh_dict
from test_gpu_parity import MESHES, setup_case

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parents[1]
CHECKER = ROOT / "tests" / "fortran" / "bin" / "check_do_oce_adv_tra"
SCHEMES = [("MUSCL", "QR4C", "FCT"), ("MFCT", "PPM", "FCT"), ("UPW1", "CDIFF", "FCT"), ("MUSCL", "UPW1", "NONE"),
           ("MFCT", "QR4C", "NONE")]


def _wet(m, d):
    L = d["nl"] - 1
    return np.arange(L)[:, None] < (m["nlevels_nod2D"] - 1)[None, :]


def _rel(a, b, mask):
    return float(np.abs(a - b)[mask].max() / max(np.abs(b)[mask].max(), 1e-300))


@pytest.mark.parametrize("mesh_name", ["periodic_24x20x12_random_bottom", "closed_16x14x10_random_bottom"])
@pytest.mark.parametrize("schemes", SCHEMES, ids=lambda s: "-".join(s))
def test_tendencies_match_oracle(api, mesh_name, schemes):
    import oracle as orc
    print(f"[gpu] do_oce_adv_tra tendencies {mesh_name} {schemes}", flush=True)
    d, dt, trs = setup_case(api, MESHES[mesh_name], schemes, ntr=3)
    m = mesh_dict(api)
    o = orc.Oracle(m)
    uv, w, we = api.mesh_array("uv"), api.mesh_array("w"), api.mesh_array("w_e")
    wet = _wet(m, d)
    rng = np.random.default_rng(3)
    for i, t in enumerate(trs):
        h, v = api.do_oce_adv_tra(i + 1, dt)
        hr, vr = o.tendencies(dt, uv, w, we, t, hor=schemes[0], ver=schemes[1], lim=schemes[2])
        eh, ev = _rel(h, hr, wet), _rel(v, vr, wet)
        assert eh <= 1e-12 and ev <= 1e-12, f"tracer {i + 1}: dttf_h {eh:.3e}, dttf_v {ev:.3e}"
        assert np.all(h[~wet] == 0.0) and np.all(v[~wet] == 0.0)
        assert np.array_equal(api.get_tracer_values(i + 1), t)          # the tracer is not touched
        # the routine adds into what the arrays hold (below the bottom they are left alone)
        h0 = np.asfortranarray(rng.normal(size=h.shape))
        v0 = np.asfortranarray(rng.normal(size=v.shape))
        h2, v2 = api.do_oce_adv_tra(i + 1, dt, h0.copy(order="F"), v0.copy(order="F"))
        assert np.array_equal(h2[~wet], h0[~wet]) and np.array_equal(v2[~wet], v0[~wet])
        assert np.allclose(h2[wet], (h0 + h)[wet], rtol=0, atol=4e-16 * (np.abs(h).max() + 4))
        assert np.allclose(v2[wet], (v0 + v)[wet], rtol=0, atol=4e-16 * (np.abs(v).max() + 4))
    # a step after the tendency calls is the ordinary step
    api.advect_steps(1, dt)
    for i, t in enumerate(trs):
        ref = np.array(t, order="F", copy=True)
        o.step(dt, uv, w, we, ref, hor=schemes[0], ver=schemes[1], lim=schemes[2])
        assert _rel(api.get_tracer_values(i + 1), ref, wet) <= 1e-12


@pytest.mark.parametrize("schemes", [("MUSCL", "QR4C", "FCT"), ("UPW1", "PPM", "NONE")], ids=lambda s: "-".join(s))
def test_accumulate_del_ttf_matches_reference_interface(api, schemes):
    import oracle as orc
    print(f"[gpu] accumulate_del_ttf {schemes}", flush=True)
    d, dt, trs = setup_case(api, MESHES["periodic_24x20x12_random_bottom"], schemes, ntr=3)
    m = mesh_dict(api)
    uv, w, we = api.mesh_array("uv"), api.mesh_array("w"), api.mesh_array("w_e")
    wet = _wet(m, d)
    api.set_option("accumulate_del_ttf", 1)
    try:
        # tracer_advect_step: tracer 1 only, del_ttf kept across calls
        o1 = orc.Oracle(m)
        ref = np.array(trs[0], order="F", copy=True)
        zeroed = np.array(trs[0], order="F", copy=True)
        oz = orc.Oracle(m)
        for step in range(3):
            assert api.advect_step(dt) == 0
            o1.step(dt, uv, w, we, ref, hor=schemes[0], ver=schemes[1], lim=schemes[2], keep_del_ttf=True)
            oz.step(dt, uv, w, we, zeroed, hor=schemes[0], ver=schemes[1], lim=schemes[2])
            err = _rel(api.get_tracer_values(1), ref, wet)
            assert err <= 1e-12, f"step {step + 1}: {err:.3e}"
        assert _rel(zeroed, ref, wet) > 1e-6                           # the two modes do differ
        # all tracers: one del_ttf per tracer
        for i, t in enumerate(trs):
            api.set_values(i + 1, t)
        api.set_option("accumulate_del_ttf", 1)                        # starts again from zero
        oracles = [orc.Oracle(m) for _ in trs]
        refs = [np.array(t, order="F", copy=True) for t in trs]
        for step in range(3):
            api.advect_steps(1, dt)
            for i in range(len(trs)):
                oracles[i].step(dt, uv, w, we, refs[i], hor=schemes[0], ver=schemes[1], lim=schemes[2],
                                keep_del_ttf=True)
                assert _rel(api.get_tracer_values(i + 1), refs[i], wet) <= 1e-12
        # graph replay of several steps keeps accumulating
        api.advect_steps(2, dt)
        for i in range(len(trs)):
            for _ in range(2):
                oracles[i].step(dt, uv, w, we, refs[i], hor=schemes[0], ver=schemes[1], lim=schemes[2],
                                keep_del_ttf=True)
            assert _rel(api.get_tracer_values(i + 1), refs[i], wet) <= 1e-12
    finally:
        api.set_option("accumulate_del_ttf", 0)


def _write_fesom_files(tmp, api, d, m, dt, trs, schemes):
    """mesh.bin / state.bin for the Fortran host: FESOM's 1-based indices, 0 for none."""
    one = lambda a: (np.asarray(a, dtype=np.int64) + 1).astype(np.int32)  # noqa: E731
    with open(tmp / "mesh.bin", "wb") as f:
        np.array([d["nl"], d["nod2D"], d["elem2D"], d["edge2D"], d["maxAdj"]], np.int32).tofile(f)
        for k in ("edges", "edge_tri", "elem2D_nodes"):
            one(m[k]).tofile(f)
        for k in ("nlevels", "ulevels", "nlevels_nod2D", "ulevels_nod2D", "nlevels_nod2D_min", "ulevels_nod2D_max",
                  "nod_in_elem2D_num"):
            np.asarray(m[k], np.int32).tofile(f)
        one(m["nod_in_elem2D"]).tofile(f)
        one(m["edge_up_dn_tri"]).tofile(f)
        np.asarray(m["nboundary_lay"], np.int32).tofile(f)
        np.asarray(api.mesh_array("zbar"), np.float64).tofile(f)
        for k in ("elem_cos", "elem_area", "gradient_sca", "edge_dxdy", "edge_cross_dxdy", "area", "areasvol",
                  "hnode", "hnode_new", "helem", "Z_3d_n", "zbar_3d_n"):
            np.asarray(m[k], np.float64).tofile(f)
    with open(tmp / "state.bin", "wb") as f:
        np.array([len(trs)], np.int32).tofile(f)
        np.array([dt], np.float64).tofile(f)
        for k in ("uv", "w", "w_e"):
            np.asarray(api.mesh_array(k), np.float64).tofile(f)
        for t, (hor, ver, lim) in zip(trs, schemes):
            np.asarray(t, np.float64).ravel(order="F").tofile(f)
            np.asarray(t, np.float64).ravel(order="F").tofile(f)          # valuesAB = values
            for sname in (hor, ver, lim):
                f.write(sname.ljust(20).encode())
            np.array([1.0, 1.0], np.float64).tofile(f)


def test_fortran_host_do_oce_adv_tra(api, tmp_path):
    import oracle as orc
    if not CHECKER.exists():
        pytest.skip("the Fortran checker was not built (no amdflang)")
    print("[gpu] Fortran host calling do_oce_adv_tra", flush=True)
    schemes = [("MUSCL", "QR4C", "FCT"), ("MFCT", "PPM", "FCT"), ("UPW1", "UPW1", "NONE")]
    d, dt, trs = setup_case(api, MESHES["periodic_24x20x12_random_bottom"], schemes[0], ntr=3)
    m = mesh_dict(api)
    _write_fesom_files(tmp_path, api, d, m, dt, trs, schemes)
    r = subprocess.run([str(CHECKER), str(tmp_path)], capture_output=True, text=True, timeout=120)
    print(r.stdout[-2000:], r.stderr[-2000:], flush=True)
    assert r.returncode == 0, r.stdout + r.stderr
    L, N = d["nl"] - 1, d["nod2D"]
    out = np.fromfile(tmp_path / "tend.bin", dtype=np.float64).reshape(len(trs), 2, 2, N, L)
    o = orc.Oracle(m)
    uv, w, we = api.mesh_array("uv"), api.mesh_array("w"), api.mesh_array("w_e")
    wet = _wet(m, d)
    for i, (t, sc) in enumerate(zip(trs, schemes)):
        hr, vr = o.tendencies(dt, uv, w, we, t, hor=sc[0], ver=sc[1], lim=sc[2])
        h1, v1 = out[i, 0, 0].T, out[i, 0, 1].T
        h2, v2 = out[i, 1, 0].T, out[i, 1, 1].T
        assert _rel(h1, hr, wet) <= 1e-12 and _rel(v1, vr, wet) <= 1e-12, f"tracer {i + 1} {sc}"
        assert np.allclose(h2, 2 * h1, rtol=1e-15, atol=0) and np.allclose(v2, 2 * v1, rtol=1e-15, atol=0)
