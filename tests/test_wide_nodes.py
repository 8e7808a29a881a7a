"""Meshes with nodes of more than 8 edges (host side, CPU only).

The kernels keep each node's edges, elements and 1-ring in a fixed-width
record: 8 wide for ordinary meshes, 16 wide when some node needs more. These
tests pin the mesh the GPU test of the 16-wide records runs on:

  * a Delaunay mesh with "wheel" hubs of 12 and 16 spokes, written as FESOM
    mesh files, loads; the hubs have exactly that many edges and elements and
    maxAdj is 16; every scheme (FCT included: a hub's elements form one fan)
    is accepted;
  * the oracle on that mesh keeps a constant constant, conserves mass to
    1e-13 and, with FCT, creates no new extrema;
  * a node with 17 edges is refused at load with an error naming the node.
"""
import numpy as np
import pytest

from conftest import mesh_dict, write_wheel_mesh


@pytest.fixture(scope="module")
def wheel(tmp_path_factory):
    d = tmp_path_factory.mktemp("wheel")
    deg = write_wheel_mesh(d, spokes=(12, 16))
    return d, deg


def test_hubs_load_with_their_degree(api, wheel):
    d, deg = wheel
    dims = api.load_mesh_from_files(d)
    assert dims["nod2D"] == len(deg) and dims["maxAdj"] == 16
    ed = api.mesh_array("edges").reshape(-1, 2)
    got = np.bincount(ed.ravel(), minlength=dims["nod2D"])
    assert np.array_equal(got, deg)
    assert sorted(deg[deg > 9].tolist()) == [12, 16]
    nie = api.mesh_array("nod_in_elem2D_num")
    hubs = np.flatnonzero(deg > 9)
    assert np.array_equal(nie[hubs], deg[hubs])  # interior hubs: one element per edge
    api.allocate_tracers(1)
    for schemes in (("MUSCL", "QR4C", "FCT"), ("MFCT", "PPM", "LEARNED"), ("UPW1", "UPW1", "NONE")):
        api.set_schemes(1, *schemes)


@pytest.mark.parametrize("schemes", [("MUSCL", "QR4C", "FCT"), ("MFCT", "PPM", "FCT"), ("UPW1", "CDIFF", "NONE")],
                         ids=lambda s: "-".join(s))
def test_oracle_step_on_wheel_mesh(api, wheel, schemes):
    import oracle as orc
    d, _ = wheel
    dims = api.load_mesh_from_files(d)
    m = mesh_dict(api)
    api.allocate_tracers(1)
    dt = api.make_velocity(0.1, 0.0, 0.2, 0.5, 0.3)
    o = orc.Oracle(m)
    uv, w, we = api.mesh_array("uv"), api.mesh_array("w"), api.mesh_array("w_e")
    L, N = dims["nl"] - 1, dims["nod2D"]
    act = np.arange(L)[:, None] < (m["nlevels_nod2D"] - 1)[None, :]
    one = np.asfortranarray(np.where(act, 1.0, 0.0))
    for _ in range(3):
        o.step(dt, uv, w, we, one, hor=schemes[0], ver=schemes[1], lim=schemes[2])
    assert np.abs(one[act] - 1.0).max() < 1e-12
    X = m["coord_nod2D"].reshape(-1, 2)
    T = np.asfortranarray(np.where(act, 10 + 5 * np.sin(40 * X[:, 0]) * np.cos(40 * X[:, 1])
                                   - 0.1 * np.arange(L)[:, None], 0))
    hn = m["hnode_new"].reshape(N, L).T
    av = m["areasvol"].reshape(N, dims["nl"]).T[:L]
    m0, lo, hi = (T * hn * av)[act].sum(), T[act].min(), T[act].max()
    T0 = T.copy()
    for _ in range(10):
        o.step(dt, uv, w, we, T, hor=schemes[0], ver=schemes[1], lim=schemes[2])
    assert abs((T * hn * av)[act].sum() - m0) <= 1e-13 * abs(m0)
    assert np.abs(T - T0).max() > 1e-3
    if schemes[2] == "FCT":
        assert T[act].min() >= lo - 1e-12 * abs(hi) and T[act].max() <= hi + 1e-12 * abs(hi)


def test_degree_17_refused(api, tmp_path):
    deg = write_wheel_mesh(tmp_path, spokes=(17,))
    hub = int(np.argmax(deg))
    assert deg[hub] == 17
    msg = api.lib.tracer_load_mesh_from_files(str(tmp_path).encode())
    assert msg == -1
    err = api.lib.tracer_last_error().decode()
    assert f"node {hub + 1} has 17 edges or elements" in err and "at most 16" in err
