"""The reference's printed numbers (README.md:51-52) from the GPU, through the C ABI.

The reference's analytic driver (src/fesom_analytic.F90) advects T on the
doubly periodic 20 x 20 x 10 analytic mesh and prints min/max/sum after each
step. Here the mesh goes in through tracer_mesh_begin / tracer_set_mesh_array
/ tracer_mesh_end exactly as test_readme_golden.py restates the driver's mesh
(halo geometry folded by the post-remap, the generator's edge handedness,
nod_in_elem2D without the elements across the seam). T is advected with
UPW1/UPW1/NONE, dt = 1000 s, u = 0.1 m/s, valuesAB set once and never
refreshed, by tracer_advect_step; tracer_get_stats, printed in the driver's
format, must equal both printed lines.

On the same mesh, given exact element gradients (the driver's are zero) and
a vertical velocity, every scheme family runs against the oracle at the FP64
tolerance of the other parity tests (1e-12 relative max-norm per step). The
seam nodes' one-sided element lists make open fans there, so FCT's cluster
bounds see the elements the reference's own loops see; edge_up_dn_tri is -1
(no up/down-wind element), so MUSCL takes its nodal-gradient branch on every
edge.
"""
import numpy as np
import pytest

from test_gpu_parity import RTOL, rel_err
from test_readme_golden import analytic_periodic_mesh, fortran_e, golden_lines, initial_temperature

pytestmark = pytest.mark.gpu

LX = LY = 1e5
NL = 10


def _gradient_sca(m):
    """Coefficients of the exact gradient of the linear interpolant on each element, (6, elem2D): the x part
    then the y part, from node differences folded to the nearest periodic image (metres)."""
    R = 6367500.0
    xy = m["coord_nod2D"].reshape(-1, 2) * R
    tri = m["elem2D_nodes"].reshape(-1, 3)

    def fold(d, L):
        return d - L * np.round(d / L)

    d2 = np.stack([fold(xy[tri[:, 1], 0] - xy[tri[:, 0], 0], LX), fold(xy[tri[:, 1], 1] - xy[tri[:, 0], 1], LY)], 1)
    d3 = np.stack([fold(xy[tri[:, 2], 0] - xy[tri[:, 0], 0], LX), fold(xy[tri[:, 2], 1] - xy[tri[:, 0], 1], LY)], 1)
    det = d2[:, 0] * d3[:, 1] - d3[:, 0] * d2[:, 1]
    gx = np.stack([d2[:, 1] - d3[:, 1], d3[:, 1], -d2[:, 1]], 1) / det[:, None]
    gy = np.stack([d3[:, 0] - d2[:, 0], -d3[:, 0], d2[:, 0]], 1) / det[:, None]
    # exact for a linear field: f = 3x - 2y sampled at the unfolded corners
    f = np.stack([np.zeros(len(tri)), 3 * d2[:, 0] - 2 * d2[:, 1], 3 * d3[:, 0] - 2 * d3[:, 1]], 1)
    np.testing.assert_allclose((gx * f).sum(1), 3.0, rtol=1e-12)
    np.testing.assert_allclose((gy * f).sum(1), -2.0, rtol=1e-12)
    return np.concatenate([gx, gy], 1).ravel()


def _load(api, gradients=False):
    m = analytic_periodic_mesh(20, 20, NL, LX, LY, 1000.0)
    if gradients:
        m["gradient_sca"] = _gradient_sca(m)
    dims = {k: m[k] for k in ("nl", "nod2D", "elem2D", "edge2D", "maxAdj")}
    arrays = {k: v for k, v in m.items() if k not in dims}
    arrays["zbar"] = m["zbar_3d_n"][:NL]
    api.set_mesh(dims["nl"], dims["nod2D"], dims["elem2D"], dims["edge2D"], dims["maxAdj"], arrays)
    api.set_option("precision", 64)
    vel = np.zeros((2, NL - 1, m["elem2D"]), order="F")
    vel[0] = 0.1
    w = np.zeros((NL, m["nod2D"]), order="F")
    return m, vel, w


def test_readme_lines_from_the_gpu(api):
    print("[gpu] README golden lines, caller-supplied analytic mesh", flush=True)
    m, vel, w = _load(api)
    api.allocate_tracers(1)
    api.set_velocity_elem(vel, w, w)
    t = initial_temperature(m["coord_nod2D"], NL, LX, LY)
    api.set_option("refresh_valuesAB", 0)
    try:
        api.set_values(1, t)                      # values and valuesAB, as the driver's copy before the loop
        api.set_schemes(1, "UPW1", "UPW1", "NONE")
        got = []
        for s in (1, 2):
            assert api.advect_step(1000.0) == 0
            st = api.get_statistics()
            v = api.get_tracer_values(1)
            got.append([str(s), fortran_e(st["min"], 14, 6).strip(), fortran_e(st["max"], 14, 6).strip(),
                        fortran_e(st["sum"], 18, 10).strip(), str(int(np.isnan(v).sum()))])
            print(f"[gpu]   {' '.join(got[-1])}", flush=True)
    finally:
        api.set_option("refresh_valuesAB", 1)
    assert got == golden_lines()[:2]


@pytest.mark.parametrize("schemes", [("UPW1", "UPW1", "NONE"), ("MUSCL", "QR4C", "FCT"), ("MFCT", "PPM", "FCT"),
                                     ("UPW1", "CDIFF", "FCT")], ids=lambda s: "-".join(s))
def test_caller_mesh_matches_oracle(api, schemes):
    import oracle as orc
    print(f"[gpu] caller-supplied analytic mesh vs oracle {schemes}", flush=True)
    m, vel, w = _load(api, gradients=True)
    vel[1] = 0.04 * np.sin(np.arange(m["elem2D"]) * 0.37)[None, :]   # a cross-flow so every edge carries flux
    w[1:NL - 1] = 2e-4 * np.cos(np.arange(m["nod2D"]) * 0.71)[None, :] * np.sin(np.arange(1, NL - 1) * 0.9)[:, None]
    api.allocate_tracers(2)
    api.set_velocity_elem(vel, w, w)
    mesh = {k: api.mesh_array(k) for k in orc.MESH_INT + orc.MESH_DBL}
    mesh.update(api.dims())
    o = orc.Oracle(mesh)
    t0 = initial_temperature(m["coord_nod2D"], NL, LX, LY)
    rng = np.random.default_rng(2)
    trs = [t0, np.asfortranarray(rng.uniform(0.0, 1.0, t0.shape))]
    ref = [np.array(t, order="F", copy=True) for t in trs]
    for i, t in enumerate(trs):
        api.set_values(i + 1, t)
        api.set_schemes(i + 1, *schemes)
    for step in range(3):
        api.advect_steps(1, 1000.0)
        for i in range(2):
            o.step(1000.0, vel, w, w, ref[i], hor=schemes[0], ver=schemes[1], lim=schemes[2])
            err = rel_err(api.get_tracer_values(i + 1), ref[i])
            assert err <= RTOL, f"tracer {i + 1} step {step + 1}: rel err {err:.3e}"
