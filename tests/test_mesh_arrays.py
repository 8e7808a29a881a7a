"""Caller-supplied meshes (tracer_mesh_begin / tracer_set_mesh_array / tracer_mesh_end), host side.

  * a generated mesh handed back with only the required arrays: every
    optional array the library derives (find_levels_min_e2n,
    nboundary_lay, Z/Z_3d_n/zbar_3d_n, depth, areasvol, hnode_new) equals
    what the generator's own derivation produced;
  * validation: a missing array, a wrong size, an index out of range, an
    element deeper than one of its nodes' columns, an unknown name and a call
    out of order are refused with an error, and no mesh is adopted;
  * the reference's analytic mesh (lib/analytic_mesh.F90, restated in
    test_readme_golden.py) is accepted as given. Its nod_in_elem2D lists, at
    the nodes on the periodic seam, only the elements on one side; those
    still form one open fan, as at a coast, so every scheme is accepted;
  * a node whose element list splits into two groups (no single fan) makes
    FCT and LEARNED refused while the unlimited schemes are accepted.
"""
import numpy as np
import pytest

from test_readme_golden import analytic_periodic_mesh

REQUIRED = ["edges", "edge_tri", "elem2D_nodes", "nlevels", "nlevels_nod2D", "nod_in_elem2D_num", "nod_in_elem2D",
            "zbar", "elem_cos", "elem_area", "gradient_sca", "edge_dxdy", "edge_cross_dxdy", "area", "hnode", "helem"]
DERIVED = ["nlevels_nod2D_min", "ulevels_nod2D_max", "nboundary_lay", "ulevels", "ulevels_nod2D", "Z", "Z_3d_n",
           "zbar_3d_n", "depth", "areasvol", "hnode_new"]


def _generated(api, periodic):
    api.generate_mesh(18, 14, 9, periodic=periodic, Lx=1.8e5, Ly=1.4e5, dz=10.0, stretch=0.3, lev_lo=3, lev_hi=9,
                      seed=3, sfc=True)
    d = api.dims()
    arrays = {k: api.mesh_array(k) for k in REQUIRED + DERIVED + ["edge_up_dn_tri", "coord_nod2D"]}
    return d, arrays


@pytest.mark.parametrize("periodic", [True, False], ids=["periodic", "closed"])
def test_required_arrays_rebuild_the_derived_ones(api, periodic):
    d, full = _generated(api, periodic)
    given = {k: full[k] for k in REQUIRED + ["edge_up_dn_tri", "coord_nod2D"]}
    got = api.set_mesh(d["nl"], d["nod2D"], d["elem2D"], d["edge2D"], d["maxAdj"], given)
    assert got == d
    for k in REQUIRED + DERIVED + ["edge_up_dn_tri", "coord_nod2D"]:
        assert np.array_equal(api.mesh_array(k), full[k]), k
    # the edge orientation the generator gives every closed mesh has open fans at the boundary: FCT allowed
    api.allocate_tracers(1)
    api.set_schemes(1, "MUSCL", "QR4C", "FCT")


def test_optional_arrays_as_given(api):
    d, full = _generated(api, True)
    given = {k: full[k] for k in REQUIRED}
    given["nboundary_lay"] = np.full(d["nod2D"], 1, np.int32)
    given["areasvol"] = 0.5 * full["area"]
    api.set_mesh(d["nl"], d["nod2D"], d["elem2D"], d["edge2D"], d["maxAdj"], given)
    assert np.array_equal(api.mesh_array("nboundary_lay"), given["nboundary_lay"])
    assert np.array_equal(api.mesh_array("areasvol"), given["areasvol"])
    assert np.all(api.mesh_array("edge_up_dn_tri") == -1)


def _rejects(api, d, arrays, match):
    with pytest.raises(RuntimeError, match=match):
        api.set_mesh(d["nl"], d["nod2D"], d["elem2D"], d["edge2D"], d["maxAdj"], arrays)
    with pytest.raises(RuntimeError, match="mesh topology missing"):
        api.allocate_tracers(1)


def test_validation(api):
    d, full = _generated(api, True)
    base = {k: full[k] for k in REQUIRED}
    _rejects(api, d, {k: v for k, v in base.items() if k != "helem"}, "helem was not set")
    _rejects(api, d, {**base, "elem_area": full["elem_area"][:-1]}, "elem_area takes")
    bad = full["edges"].copy()
    bad[7] = d["nod2D"]
    _rejects(api, d, {**base, "edges": bad}, "edges: bad entry at 7")
    bad = full["nod_in_elem2D"].copy()
    bad[0] = -1
    _rejects(api, d, {**base, "nod_in_elem2D": bad}, "nod_in_elem2D")
    bad = full["nlevels"].copy()
    el = int(np.argmin(bad))                    # a shallow element made deeper than its nodes' columns
    bad[el] = d["nl"]
    assert np.any(full["nlevels_nod2D"][full["elem2D_nodes"].reshape(-1, 3)[el]] < d["nl"])
    _rejects(api, d, {**base, "nlevels": bad}, "nlevels_nod2D")
    _rejects(api, d, {**base, "no_such_array": np.zeros(3)}, "unknown mesh array")
    lib = api.lib
    assert lib.tracer_mesh_end() != 0           # no begin since the last failure
    assert lib.tracer_set_mesh_array(b"zbar", None, 0) != 0
    assert lib.tracer_mesh_begin(2, 10, 10, 10, 6) != 0


def test_reference_analytic_mesh_accepted(api):
    m = analytic_periodic_mesh(20, 20, 10, 1e5, 1e5, 1000.0)
    dims = {k: m[k] for k in ("nl", "nod2D", "elem2D", "edge2D", "maxAdj")}
    arrays = {k: v for k, v in m.items() if k not in dims}
    arrays["zbar"] = m["zbar_3d_n"][:m["nl"]]
    got = api.set_mesh(dims["nl"], dims["nod2D"], dims["elem2D"], dims["edge2D"], dims["maxAdj"], arrays)
    assert got == dims
    for k, v in arrays.items():
        assert np.array_equal(api.mesh_array(k), np.ravel(v)), k
    assert np.min(m["nod_in_elem2D_num"]) < 6     # the seam nodes' short lists
    api.allocate_tracers(2)
    api.set_schemes(1, "UPW1", "UPW1", "FCT")
    api.set_schemes(2, "MUSCL", "QR4C", "LEARNED")


def test_split_element_list_refuses_fct(api):
    d, full = _generated(api, True)
    A = d["maxAdj"]
    nie = full["nod_in_elem2D"].reshape(-1, A).copy()
    num = full["nod_in_elem2D_num"].copy()
    tri = full["elem2D_nodes"].reshape(-1, 3)
    n = int(np.argmax(num == 6))
    els = list(nie[n, :6])
    others = [set(tri[e]) - {n} for e in els]
    # two elements opposite in the fan (no shared ring node): removing both leaves two groups
    i, j = next((i, j) for i in range(6) for j in range(i + 1, 6) if not others[i] & others[j])
    keep = [e for k, e in enumerate(els) if k not in (i, j)]
    nie[n, :] = 0
    nie[n, :4] = keep
    num[n] = 4
    given = {k: full[k] for k in REQUIRED}
    given.update(nod_in_elem2D=nie.ravel(), nod_in_elem2D_num=num)
    api.set_mesh(d["nl"], d["nod2D"], d["elem2D"], d["edge2D"], A, given)
    api.allocate_tracers(2)
    api.set_schemes(1, "UPW1", "UPW1", "NONE")
    api.set_schemes(2, "MUSCL", "QR4C", "NONE")
    with pytest.raises(RuntimeError, match=f"one fan; node {n + 1} does not"):
        api.set_schemes(1, "UPW1", "UPW1", "FCT")
    with pytest.raises(RuntimeError, match="one fan"):
        api.set_schemes(2, "MUSCL", "QR4C", "LEARNED")
