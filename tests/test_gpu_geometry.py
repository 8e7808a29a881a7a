"""Mesh geometry on the GPU (option "device_geometry": mesh_areas and mesh_auxiliary_arrays as kernels).

  * planar meshes (periodic and closed, random bottoms): every array the
    setup derives (elem_area, elem_cos, gradient_sca, edge_dxdy,
    edge_cross_dxdy, area, areasvol, and what the host derives from them)
    equals the host derivation bit for bit;
  * the reference's pi mesh on the sphere: integer arrays equal, floating
    point ones within 1e-14 relative (the device cos may differ in the last
    bit), and one MUSCL+QR4C+FCT step on either setup agrees to the FP64
    parity tolerance.
"""
import time

import numpy as np
import pytest

import oracle as orc

pytestmark = pytest.mark.gpu

EXTRA = ["elem_cos", "elem_area", "gradient_sca", "edge_dxdy", "edge_cross_dxdy", "area", "areasvol",
         "edge_up_dn_tri", "nboundary_lay", "nlevels_nod2D_min", "depth", "hnode", "Z_3d_n", "zbar_3d_n"]


def _arrays(api):
    return {k: api.mesh_array(k) for k in sorted(set(orc.MESH_INT) | set(orc.MESH_DBL) | set(EXTRA))}


@pytest.mark.parametrize("periodic", [True, False], ids=["periodic", "closed"])
def test_planar_geometry_bitwise(api, periodic):
    print(f"[gpu] device geometry, planar periodic={periodic}", flush=True)
    kw = dict(periodic=periodic, Lx=3.1e5, Ly=2.3e5, dz=10.0, stretch=0.4, lev_lo=3, lev_hi=20, seed=11, sfc=True)
    try:
        api.set_option("device_geometry", 0)
        t = time.perf_counter()
        api.generate_mesh(310, 230, 20, **kw)
        t_host = time.perf_counter() - t
        host = _arrays(api)
        api.set_option("device_geometry", 1)
        t = time.perf_counter()
        api.generate_mesh(310, 230, 20, **kw)
        t_dev = time.perf_counter() - t
        dev = _arrays(api)
    finally:
        api.set_option("device_geometry", 0)
    print(f"[gpu]   generate_mesh 310x230x20: host {t_host:.2f} s, device geometry {t_dev:.2f} s", flush=True)
    for k in host:
        assert np.array_equal(host[k], dev[k]), k
    assert np.any(host["edge_up_dn_tri"] >= 0)


def test_sphere_geometry_and_step(api, pi_mesh_dir):
    print("[gpu] device geometry, pi mesh", flush=True)
    from test_gpu_parity import RTOL, rel_err
    res = {}
    try:
        for flag in (0, 1):
            api.set_option("device_geometry", flag)
            api.load_mesh_from_files(pi_mesh_dir)
            arrays = _arrays(api)
            api.allocate_tracers(1)
            dt = api.make_velocity(0.1, 0.0, 0.2, 0.5, 0.3)
            lat = arrays["coord_nod2D"].reshape(-1, 2)[:, 1]
            nl = api.dims()["nl"]
            t0 = np.asfortranarray(np.tile(10.0 + 5.0 * np.sin(3.0 * lat), (nl - 1, 1))
                                   + np.linspace(0.0, 2.0, nl - 1)[:, None])
            api.set_values(1, t0)
            api.set_schemes(1, "MUSCL", "QR4C", "FCT")
            api.advect_steps(2, dt)
            res[flag] = (arrays, api.get_tracer_values(1), dt)
    finally:
        api.set_option("device_geometry", 0)
    host, dev = res[0][0], res[1][0]
    for k in host:
        if host[k].dtype.kind == "i":
            assert np.array_equal(host[k], dev[k]), k
        else:
            np.testing.assert_allclose(dev[k], host[k], rtol=1e-14, atol=1e-14 * np.abs(host[k]).max(), err_msg=k)
    assert abs(res[1][2] - res[0][2]) <= 1e-12 * res[0][2]
    err = rel_err(res[1][1], res[0][1])
    print(f"[gpu]   two steps, device vs host geometry: rel err {err:.2e}", flush=True)
    assert err <= RTOL
