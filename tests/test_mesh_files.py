"""FESOM mesh-file reader (tracer_load_mesh_from_files) on the reference's own
pi mesh (tests/data/pi: 3140 nodes, 5839 triangles, 8986 edges, 48 levels),
and the spherical geometry derived from it. CPU only.

What pins the reader and the geometry:
  * the files' own cross-references: counts, edgenum.out's inner-edge count
    equals the edges with one element, nlvls.out equals the deepest level of
    each node's elements, every edge's nodes belong to its elements;
  * the rotation (FESOM's Euler angles 50, 15, -90 degrees) is an isometry
    that sends the geographic point (-40 E, 75 N) to the rotated pole and the
    geographic pole to rotated (180 E, 75 N) -- closed forms of the angles;
  * orientation: every element clockwise and edge_tri(1) left of its edge in
    the rotated frame (FESOM's convention, lib/MOD_MESH.F90);
  * gradient_sca is the exact gradient of fields linear in each element's
    local frame (x = r cos(lat_c) lon, y = r lat);
  * elem_area against the exact spherical-excess area of each triangle
    (l'Huilier), computed from the geographic file coordinates alone;
  * the whole step on this mesh (oracle): a constant stays constant, mass is
    conserved and FCT creates no new extrema.
"""
import ctypes

import numpy as np
import pytest

from conftest import mesh_dict

R = 6367500.0
RAD = np.pi / 180.0


def _read(d, name, dtype=float):
    return np.array((d / name).read_text().split(), dtype=dtype)


def _geo(d):
    t = (d / "nod2d.out").read_text().split()
    n = int(t[0])
    a = np.array(t[1:], dtype=float).reshape(n, 4)
    return a[:, 1] * RAD, a[:, 2] * RAD


def _unit(lon, lat):
    return np.stack([np.cos(lat) * np.cos(lon), np.cos(lat) * np.sin(lon), np.sin(lat)], axis=-1)


@pytest.fixture(scope="module")
def pi(api, pi_mesh_dir):
    api.set_option("force_rotation", 1)
    d = api.load_mesh_from_files(pi_mesh_dir)
    return d, mesh_dict(api)


def test_counts_and_cross_references(pi, pi_mesh_dir):
    d, m = pi
    assert (d["nod2D"], d["elem2D"], d["edge2D"], d["nl"]) == (3140, 5839, 8986, 48)
    edge2D, edge2D_in = _read(pi_mesh_dir, "edgenum.out", int)
    et = m["edge_tri"].reshape(-1, 2)
    assert (et[:, 0] >= 0).all() and (et[:, 1] < 0).sum() == edge2D - edge2D_in == 455
    assert np.array_equal(m["nlevels"], _read(pi_mesh_dir, "elvls.out", int))
    assert np.array_equal(m["nlevels_nod2D"], _read(pi_mesh_dir, "nlvls.out", int))
    en = m["elem2D_nodes"].reshape(-1, 3)
    ed = m["edges"].reshape(-1, 2)
    for s in (0, 1):
        has = et[:, s] >= 0
        tri = en[et[has, s]]
        for q in (0, 1):
            assert (tri == ed[has, q][:, None]).any(axis=1).all()
    zbar = _read(pi_mesh_dir, "aux3d.out")[1:49]
    assert np.array_equal(api_zbar(m), -np.abs(zbar))


def api_zbar(m):
    # zbar_3d_n of any node is the standard zbar column
    return m["zbar_3d_n"].reshape(-1, 48)[0]


def test_rotation_is_fesoms_euler_rotation(pi, pi_mesh_dir):
    _, m = pi
    glon, glat = _geo(pi_mesh_dir)
    r = m["coord_nod2D"].reshape(-1, 2)
    g3, r3 = _unit(glon, glat), _unit(r[:, 0], r[:, 1])
    # isometry: angular distances between all pairs of a node sample are kept
    idx = np.random.default_rng(0).choice(len(glon), 200, replace=False)
    assert np.allclose(g3[idx] @ g3[idx].T, r3[idx] @ r3[idx].T, rtol=0, atol=1e-13)
    # the rotated pole is geographic (-40 E, 75 N): rotated latitude = angle to it
    pole = _unit(-40.0 * RAD, 75.0 * RAD)
    assert np.allclose(np.sin(r[:, 1]), g3 @ pole, rtol=0, atol=1e-13)
    # the geographic pole sits at rotated (180 E, 75 N)
    npole = _unit(np.pi, 75.0 * RAD)
    assert np.allclose(g3[:, 2], r3 @ npole, rtol=0, atol=1e-13)
    # orientation preserved (a proper rotation, not a reflection)
    en = m["elem2D_nodes"].reshape(-1, 3)
    vol_g = np.einsum("ij,ij->i", g3[en[:, 0]], np.cross(g3[en[:, 1]], g3[en[:, 2]]))
    vol_r = np.einsum("ij,ij->i", r3[en[:, 0]], np.cross(r3[en[:, 1]], r3[en[:, 2]]))
    assert np.allclose(vol_g, vol_r, rtol=1e-9, atol=1e-15)


def _trim(x):
    return np.where(x > np.pi, x - 2 * np.pi, np.where(x < -np.pi, x + 2 * np.pi, x))


def test_orientation_clockwise_and_edge_tri_left(pi):
    _, m = pi
    X = m["coord_nod2D"].reshape(-1, 2)
    en = m["elem2D_nodes"].reshape(-1, 3)
    b = np.stack([_trim(X[en[:, 1], 0] - X[en[:, 0], 0]), X[en[:, 1], 1] - X[en[:, 0], 1]], 1)
    c = np.stack([_trim(X[en[:, 2], 0] - X[en[:, 0], 0]), X[en[:, 2], 1] - X[en[:, 0], 1]], 1)
    assert (b[:, 0] * c[:, 1] - b[:, 1] * c[:, 0] < 0).all()
    # edge_cross_dxdy: element 1 left of the edge, element 2 right, in the local frame
    ed = m["edges"].reshape(-1, 2)
    et = m["edge_tri"].reshape(-1, 2)
    dxdy = m["edge_dxdy"].reshape(-1, 2)
    cos1 = m["elem_cos"][et[:, 0]]
    ex, ey = dxdy[:, 0] * cos1, dxdy[:, 1]
    cd = m["edge_cross_dxdy"].reshape(-1, 4)
    assert (ex * cd[:, 1] - ey * cd[:, 0] > 0).all()
    inner = et[:, 1] >= 0
    cos2 = m["elem_cos"][et[inner, 1]]
    assert (dxdy[inner, 0] * cos2 * cd[inner, 3] - dxdy[inner, 1] * cd[inner, 2] < 0).all()
    assert (cd[~inner, 2:] == 0).all()
    assert np.array_equal(dxdy[:, 0], _trim(X[ed[:, 1], 0] - X[ed[:, 0], 0]))


def test_gradient_exact_in_each_elements_local_frame(pi):
    _, m = pi
    X = m["coord_nod2D"].reshape(-1, 2)
    en = m["elem2D_nodes"].reshape(-1, 3)
    g = m["gradient_sca"].reshape(-1, 6)
    ec = m["elem_cos"]
    assert np.allclose(ec, np.cos(X[en, 1].mean(axis=1)), rtol=1e-14, atol=0)
    rng = np.random.default_rng(1)
    a, b = rng.uniform(-1, 1, 2)
    x = R * ec[:, None] * _trim(X[en, 0] - X[en[:, :1], 0])
    y = R * (X[en, 1] - X[en[:, :1], 1])
    f = a * x + b * y + 3.0
    gx = (g[:, 0:3] * f).sum(axis=1)
    gy = (g[:, 3:6] * f).sum(axis=1)
    assert np.allclose(gx, a, rtol=1e-9, atol=1e-12) and np.allclose(gy, b, rtol=1e-9, atol=1e-12)


def test_areas_against_spherical_excess(pi, pi_mesh_dir):
    _, m = pi
    glon, glat = _geo(pi_mesh_dir)
    p = _unit(glon, glat)
    en = m["elem2D_nodes"].reshape(-1, 3)
    A, B, C = p[en[:, 0]], p[en[:, 1]], p[en[:, 2]]
    num = np.abs(np.einsum("ij,ij->i", A, np.cross(B, C)))
    den = 1 + np.einsum("ij,ij->i", A, B) + np.einsum("ij,ij->i", B, C) + np.einsum("ij,ij->i", C, A)
    exact = 2 * np.arctan2(num, den) * R * R
    rel = m["elem_area"] / exact - 1
    # planar cos(lat) triangles: the total agrees to 1e-4; single triangles
    # near the rotated poles differ by up to ~9 %
    assert abs(m["elem_area"].sum() / exact.sum() - 1) < 1e-3
    assert np.percentile(np.abs(rel), 99) < 3.5e-2 and np.abs(rel).max() < 0.1
    # nodal areas: a third of every element around the node, on each of its wet levels
    area = m["area"].reshape(-1, 48)
    assert np.isclose(area[:, 0].sum(), m["elem_area"].sum(), rtol=1e-13)
    assert (area[np.arange(48)[None, :] < m["nlevels_nod2D"][:, None] - 1] > 0).all()


@pytest.mark.parametrize("schemes", [("UPW1", "UPW1", "NONE"), ("MUSCL", "QR4C", "FCT"), ("MFCT", "PPM", "FCT")],
                         ids=lambda s: "-".join(s))
def test_oracle_step_on_pi_mesh(api, pi, schemes):
    import oracle as orc
    d, m = pi
    api.allocate_tracers(1)
    dt = api.make_velocity(0.1, 0.0, 0.2, 0.5, 0.3)
    o = orc.Oracle(m)
    uv, w, we = api.mesh_array("uv"), api.mesh_array("w"), api.mesh_array("w_e")
    L, N = d["nl"] - 1, d["nod2D"]
    act = np.arange(L)[:, None] < (m["nlevels_nod2D"] - 1)[None, :]
    one = np.asfortranarray(np.where(act, 1.0, 0.0))
    for _ in range(3):
        o.step(dt, uv, w, we, one, hor=schemes[0], ver=schemes[1], lim=schemes[2])
    assert np.abs(one[act] - 1.0).max() < 1e-12
    X = m["coord_nod2D"].reshape(-1, 2)
    T = np.asfortranarray(np.where(act, 10 + 5 * np.sin(X[:, 0]) * np.cos(X[:, 1]) - 0.1 * np.arange(L)[:, None], 0))
    hn = m["hnode_new"].reshape(N, L).T
    av = m["areasvol"].reshape(N, d["nl"]).T[:L]
    m0, lo, hi = (T * hn * av)[act].sum(), T[act].min(), T[act].max()
    T0 = T.copy()
    for _ in range(10):
        o.step(dt, uv, w, we, T, hor=schemes[0], ver=schemes[1], lim=schemes[2])
    assert abs((T * hn * av)[act].sum() - m0) <= 1e-13 * abs(m0)
    assert np.abs(T - T0).max() > 1e-3
    if schemes[2] == "FCT":
        assert T[act].min() >= lo - 1e-12 * (hi - lo) and T[act].max() <= hi + 1e-12 * (hi - lo)


def test_without_rotation_and_errors(api, pi_mesh_dir, tmp_path):
    api.set_option("force_rotation", 0)
    try:
        api.load_mesh_from_files(pi_mesh_dir)
        glon, glat = _geo(pi_mesh_dir)
        X = api.mesh_array("coord_nod2D").reshape(-1, 2)
        assert np.array_equal(X[:, 0], glon) and np.array_equal(X[:, 1], glat)
    finally:
        api.set_option("force_rotation", 1)
    lib = api.lib
    assert lib.tracer_load_mesh_from_files(str(tmp_path / "missing").encode()) == -1
    assert lib.tracer_load_mesh_from_files(str(tmp_path).encode()) == -1  # no nod2d.out
    api.load_mesh_from_files(pi_mesh_dir)


def _info(path):
    out = {}
    for line in path.read_text().splitlines():
        k, v = line.split("=", 1)
        out[k] = v
    return out


def test_mesh_and_scalar_output_format(api, pi, pi_mesh_dir, tmp_path):
    """The files of lib/mesh_output.F90, read the way visualization/plot_scalars.py
    reads them (np.fromfile, Fortran order)."""
    d, m = pi
    api.load_mesh_from_files(pi_mesh_dir)
    out = tmp_path / "out"
    api.write_mesh(out)
    info = _info(out / "mesh_info.txt")
    assert (int(info["nod2D"]), int(info["elem2D"]), int(info["nl"])) == (3140, 5839, 48)
    xy = np.fromfile(out / "coord_nod2D.bin", dtype=np.float64).reshape((2, 3140), order="F")
    assert np.array_equal(xy, (api.mesh_array("coord_nod2D") * R).reshape(-1, 2).T)
    tri = np.fromfile(out / "elem2D_nodes.bin", dtype=np.int32).reshape((3, 5839), order="F").T
    assert np.array_equal(tri, api.mesh_array("elem2D_nodes").reshape(-1, 3))
    assert np.array_equal(np.fromfile(out / "zbar.bin"), api.mesh_array("zbar"))
    depth = np.fromfile(out / "depth.bin")
    raw = _read(pi_mesh_dir, "aux3d.out")[49:]
    assert np.array_equal(depth, np.minimum(-np.abs(raw), api.mesh_array("zbar")[4]))
    api.allocate_tracers(3)
    L = d["nl"] - 1
    vals = [np.asfortranarray(np.random.default_rng(t).uniform(0, 1, (L, 3140))) for t in range(3)]
    for t, v in enumerate(vals):
        api.set_values(t + 1, v)
    api.output_open(out, 1)
    api.output_step()
    api.output_step()
    api.output_close()
    sinfo = _info(out / "scalar_info.txt")
    assert sinfo["nz"] == str(L) and sinfo["nsteps"] == "1" and sinfo["wp"] == "8"
    names = [sinfo[f"tracer_{t + 1}"] for t in range(3)]
    assert names == ["temperature", "salinity", "tracer_3"]
    for t, name in enumerate(names):
        raw = np.fromfile(out / f"{name}.bin", dtype=np.float64)
        frames = raw.reshape((2, 3140, L))  # plot_scalars.py: (nframes, nod2D, nz)
        for f in range(2):
            assert np.array_equal(frames[f], vals[t].T)
