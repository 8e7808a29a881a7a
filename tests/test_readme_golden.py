"""The reference's own golden output (README.md:51-52) reproduced by the oracle.

The reference's analytic driver (src/fesom_analytic.F90) advects T on the
doubly-periodic 20 x 20 x 10 analytic mesh and prints min/max/sum after each
step. Those two printed lines are the only known-answer vectors the reference
ships for this path. To reproduce them the oracle has to run on the mesh the
driver actually builds, so the builder below restates, in numpy,
lib/analytic_mesh.F90 `generate_analytic_mesh` (periodic branch) together
with the geometry it calls (lib/oce_mesh.F90 mesh_areas and
mesh_auxiliary_arrays, edge_center, elem_center) and its post-remap, with
every property of that mesh that changes the numbers:

  * halo copies of the x = Lx column, the y = Ly row and the corner carry the
    geometry until the post-remap folds them onto real nodes
    (analytic_mesh.F90:735-827);
  * edge_tri(1) is the element the generator calls "upper", which is the
    right-hand element of horizontal and vertical edges
    (analytic_mesh.F90:333-384), so those edges carry edge_cross_dxdy of the
    opposite handedness to the rest of FESOM;
  * cross distances of wrap-around edges are folded by the min-image rule
    after they were computed on the unfolded halo mesh
    (analytic_mesh.F90:761-777);
  * valuesAB is copied once before the loop (fesom_analytic.F90:333-336)
    and the unlimited UPW1 flux reads valuesAB (oce_adv_tra_driver.F90:251),
    so step 2 sees the initial field again.

The library's own mesh generator builds a correctly oriented mesh instead;
this file exists to pin the oracle to the reference's printed numbers.
"""
from pathlib import Path

import numpy as np
import pytest

import oracle as orc

R_EARTH = 6367500.0        # lib/oce_modules.F90:24
PI = 3.14159265358979      # lib/oce_modules.F90:19
GOLDEN = Path(__file__).resolve().parent / "golden" / "readme_analytic_periodic_20x20x10.txt"


def analytic_periodic_mesh(nx, ny, nl, Lx, Ly, max_depth):
    """Mesh arrays (oracle layout: 0-based, Fortran order flattened) of the periodic analytic mesh."""
    nxr, nyr = nx - 1, ny - 1
    nloc = nxr * nyr
    edim = nxr + nyr + 1
    ntot = nloc + edim

    def pnode(i, j):  # 1-based grid point -> 0-based node (halo indices >= nloc); analytic_mesh.F90:721-733
        if i <= nxr and j <= nyr:
            return (j - 1) * nxr + i - 1
        if i == nx and j <= nyr:
            return nloc + j - 1
        if i <= nxr and j == ny:
            return nloc + nyr + i - 1
        return ntot - 1

    dx, dy = Lx / nxr, Ly / nyr
    coord = np.zeros((ntot, 2))                      # analytic_mesh.F90:263-285
    for j in range(1, ny):
        for i in range(1, nx):
            coord[pnode(i, j)] = ((i - 1) * dx / R_EARTH, (j - 1) * dy / R_EARTH)
    for j in range(1, ny):
        coord[nloc + j - 1] = (Lx / R_EARTH, (j - 1) * dy / R_EARTH)
    for i in range(1, nx):
        coord[nloc + nyr + i - 1] = ((i - 1) * dx / R_EARTH, Ly / R_EARTH)
    coord[ntot - 1] = (Lx / R_EARTH, Ly / R_EARTH)

    tri = []                                         # analytic_mesh.F90:312-332
    for j in range(1, ny):
        for i in range(1, nx):
            n1, n2, n3, n4 = pnode(i, j), pnode(i + 1, j), pnode(i + 1, j + 1), pnode(i, j + 1)
            tri += [(n1, n3, n2), (n1, n4, n3)]
    tri = np.array(tri)
    T = len(tri)

    def lower(i, j):   # element 2*((j-1)*(nx-1)+(i-1))+1, 0-based
        return 2 * ((j - 1) * nxr + (i - 1))

    edges, etri = [], []                             # analytic_mesh.F90:383-450
    for j in range(1, ny):
        for i in range(1, nx):
            up = lower(i, j - 1) + 1 if j >= 2 else lower(i, ny - 1) + 1
            edges.append((pnode(i, j), pnode(i + 1, j)))
            etri.append((up, lower(i, j)))
    for j in range(1, ny):
        for i in range(1, nx):
            lo = lower(i - 1, j) if i >= 2 else lower(nx - 1, j)
            edges.append((pnode(i, j), pnode(i, j + 1)))
            etri.append((lower(i, j) + 1, lo))
    for j in range(1, ny):
        for i in range(1, nx):
            edges.append((pnode(i, j), pnode(i + 1, j + 1)))
            etri.append((lower(i, j) + 1, lower(i, j)))
    edges, etri = np.array(edges), np.array(etri)
    E = len(edges)

    # nod_in_elem2D over the pre-remap (halo) indices, analytic_mesh.F90:667-683
    nie = [[] for _ in range(ntot)]
    for e in range(T):
        for q in range(3):
            nie[tri[e, q]].append(e)
    max_adj = max(len(v) for v in nie)

    # mesh_areas (oce_mesh.F90:2089-2204), cartesian: no cos factor, no cyclic trim at these sizes
    a = coord[tri[:, 1]] - coord[tri[:, 0]]
    b = coord[tri[:, 2]] - coord[tri[:, 0]]
    elem_area = 0.5 * np.abs(a[:, 0] * b[:, 1] - b[:, 0] * a[:, 1])
    area = np.zeros((ntot, nl))
    for n in range(ntot):
        acc = 0.0
        for e in nie[n]:
            acc = acc + elem_area[e] / 3.0
        area[n, :nl - 1] = acc
    elem_area = elem_area * R_EARTH * R_EARTH
    area = area * R_EARTH * R_EARTH
    areasvol = area.copy()

    # mesh_auxiliary_arrays (oce_mesh.F90:2397-2465): centres on the unfolded mesh
    cx = (coord[tri[:, 0], 0] + coord[tri[:, 1], 0] + coord[tri[:, 2], 0]) / 3.0
    cy = (coord[tri[:, 0], 1] + coord[tri[:, 1], 1] + coord[tri[:, 2], 1]) / 3.0
    ex = 0.5 * (coord[edges[:, 0], 0] + coord[edges[:, 1], 0])
    ey = 0.5 * (coord[edges[:, 0], 1] + coord[edges[:, 1], 1])
    edge_dxdy = coord[edges[:, 1]] - coord[edges[:, 0]]
    cross = np.zeros((E, 4))
    for s in range(2):
        cross[:, 2 * s] = (cx[etri[:, s]] - ex) * R_EARTH
        cross[:, 2 * s + 1] = (cy[etri[:, s]] - ey) * R_EARTH

    # periodic_post_remap (analytic_mesh.F90:736-827)
    for k, L in ((0, Lx), (1, Ly), (2, Lx), (3, Ly)):
        v = cross[:, k]
        fold = np.abs(v) > 0.5 * L
        cross[fold, k] = v[fold] - np.copysign(L, v[fold])
    halo_to_real = [(h - 1) * nxr for h in range(1, ny)] + list(range(nxr)) + [0]
    for h, r in enumerate(halo_to_real):
        area[r, :nl - 1] += area[nloc + h, :nl - 1]
        areasvol[r, :nl - 1] += areasvol[nloc + h, :nl - 1]

    def remap(x):
        x = x.copy()
        halo = x >= nloc
        x[halo] = np.array(halo_to_real)[x[halo] - nloc]
        return x

    tri, edges = remap(tri), remap(edges)

    zbar = np.array([-((k - 1) * max_depth / (nl - 1)) for k in range(1, nl + 1)])
    Z = 0.5 * (zbar[:-1] + zbar[1:])
    hn = zbar[:-1] - zbar[1:]                        # fesom_analytic.F90:226-238
    nie_arr = np.zeros((nloc, max_adj), dtype=np.int32)
    nie_num = np.zeros(nloc, dtype=np.int32)
    for n in range(nloc):
        nie_num[n] = len(nie[n])
        nie_arr[n, :len(nie[n])] = nie[n]
    full = lambda v, cnt: np.full(cnt, v, dtype=np.int32)  # noqa: E731
    return {
        "nl": nl, "nod2D": nloc, "elem2D": T, "edge2D": E, "maxAdj": max_adj,
        "edges": edges.ravel(), "edge_tri": etri.ravel(), "elem2D_nodes": tri.ravel(),
        "nlevels": full(nl, T), "ulevels": full(1, T), "nlevels_nod2D": full(nl, nloc),
        "ulevels_nod2D": full(1, nloc), "nlevels_nod2D_min": full(nl, nloc), "ulevels_nod2D_max": full(1, nloc),
        "nod_in_elem2D_num": nie_num, "nod_in_elem2D": nie_arr.ravel(),
        "edge_up_dn_tri": full(-1, 2 * E), "nboundary_lay": full(nl - 1, nloc),
        "coord_nod2D": coord[:nloc].ravel(), "edge_dxdy": edge_dxdy.ravel(), "edge_cross_dxdy": cross.ravel(),
        "elem_cos": np.ones(T), "elem_area": elem_area, "gradient_sca": np.zeros(6 * T),
        "area": area[:nloc].ravel(), "areasvol": areasvol[:nloc].ravel(),
        "hnode": np.tile(hn, nloc), "hnode_new": np.tile(hn, nloc), "helem": np.tile(hn, T),
        "Z_3d_n": np.tile(Z, nloc), "zbar_3d_n": np.tile(zbar, nloc),
    }


def initial_temperature(coord, nl, Lx, Ly):
    """fesom_analytic.F90:307-318: warm ridge plus linear stratification, (nl-1, nod2D) F-order."""
    nz = nl - 1
    xy = coord.reshape(-1, 2) * R_EARTH
    ridge = 5.0 * np.sin(2.0 * PI * xy[:, 0] / Lx) * np.cos(2.0 * PI * xy[:, 1] / Ly)
    zn = np.arange(nz) / (nz - 1)
    return np.asfortranarray((15.0 + ridge)[None, :] - 10.0 * zn[:, None])


def fortran_e(x, w, d):
    """Fortran Ew.d edit descriptor (gfortran): 0.ddddddE+xx, right-justified in w columns."""
    if x == 0.0:
        return ("0." + "0" * d + "E+00").rjust(w)
    m, e = f"{abs(x):.{d - 1}e}".split("e")
    digits = m.replace(".", "")
    s = ("-" if x < 0 else "") + "0." + digits + "E" + f"{int(e) + 1:+03d}"
    return s.rjust(w)


def golden_lines():
    rows = []
    for line in GOLDEN.read_text().splitlines():
        if line and not line.startswith("#"):
            rows.append(line.split())
    return rows


def test_fortran_e_format():
    assert fortran_e(0.0246394, 14, 6).strip() == "0.246394E-01"
    assert fortran_e(32490.0, 18, 10).strip() == "0.3249000000E+05"
    assert fortran_e(19.97364, 14, 6).strip() == "0.199736E+02"


def test_mesh_restatement_invariants():
    m = analytic_periodic_mesh(20, 20, 10, 1e5, 1e5, 1000.0)
    assert (m["nod2D"], m["elem2D"], m["edge2D"]) == (361, 722, 1083)
    # every real node's merged area is the full hexagon 6 * (dx*dy/2) / 3 = dx*dy
    dx = 1e5 / 19
    area = m["area"].reshape(-1, 10)
    np.testing.assert_allclose(area[:, :9], dx * dx, rtol=1e-12)
    assert np.all(area[:, 9] == 0.0)
    # every dual segment spans one element centre to the edge midpoint: |d| < dx
    cross = m["edge_cross_dxdy"].reshape(-1, 4)
    assert np.abs(cross).max() < dx
    # horizontal and vertical edges carry the generator's flipped handedness, diagonals FESOM's:
    # d(el1) x edge direction has opposite sign on the two groups
    ed = m["edge_dxdy"].reshape(-1, 2)
    z = cross[:, 0] * ed[:, 1] - cross[:, 1] * ed[:, 0]
    n = 361
    assert np.all(np.sign(z[:2 * n]) == -np.sign(z[2 * n:3 * n][0]))
    assert np.all(np.sign(z[2 * n:]) == np.sign(z[2 * n]))


@pytest.mark.parametrize("steps", [2])
def test_readme_golden_vectors(steps):
    Lx = Ly = 1e5
    nl = 10
    m = analytic_periodic_mesh(20, 20, nl, Lx, Ly, 1000.0)
    o = orc.Oracle(m)
    T, N = m["elem2D"], m["nod2D"]
    vel = np.zeros((2, nl - 1, T), order="F")
    vel[0] = 0.1
    w = np.zeros((nl, N), order="F")
    t = initial_temperature(m["coord_nod2D"], nl, Lx, Ly)
    ab = t.copy(order="F")                           # set once, never refreshed
    got = []
    for s in range(1, steps + 1):
        o.step(1000.0, vel, w, w, t, valuesAB=ab, hor="UPW1", ver="UPW1", lim="NONE")
        got.append([str(s), fortran_e(t.min(), 14, 6).strip(), fortran_e(t.max(), 14, 6).strip(),
                    fortran_e(float(np.sum(t)), 18, 10).strip(), str(int(np.isnan(t).sum()))])
    assert got == golden_lines()[:steps]


def test_readme_golden_needs_stale_ab():
    """With valuesAB refreshed every step, step 2 departs from the printed line (the staleness is load-bearing)."""
    Lx = Ly = 1e5
    nl = 10
    m = analytic_periodic_mesh(20, 20, nl, Lx, Ly, 1000.0)
    o = orc.Oracle(m)
    vel = np.zeros((2, nl - 1, m["elem2D"]), order="F")
    vel[0] = 0.1
    w = np.zeros((nl, m["nod2D"]), order="F")
    t = initial_temperature(m["coord_nod2D"], nl, Lx, Ly)
    for _ in range(2):
        o.step(1000.0, vel, w, w, t, valuesAB=None, hor="UPW1", ver="UPW1", lim="NONE")
    assert fortran_e(t.min(), 14, 6).strip() != golden_lines()[1][1]


def test_readme_golden_detects_edge_orientation():
    """Giving horizontal and vertical edges FESOM's own handedness (swap edge_tri and the two halves of
    edge_cross_dxdy) moves step 1 off the printed line: the vectors pin the sign of the dual-segment flux."""
    Lx = Ly = 1e5
    nl = 10
    m = analytic_periodic_mesh(20, 20, nl, Lx, Ly, 1000.0)
    n = m["nod2D"]
    et = m["edge_tri"].reshape(-1, 2).copy()
    cr = m["edge_cross_dxdy"].reshape(-1, 4).copy()
    et[:2 * n] = et[:2 * n, ::-1]
    cr[:2 * n] = cr[:2 * n][:, [2, 3, 0, 1]]
    m.update(edge_tri=et.ravel(), edge_cross_dxdy=cr.ravel())
    o = orc.Oracle(m)
    vel = np.zeros((2, nl - 1, m["elem2D"]), order="F")
    vel[0] = 0.1
    w = np.zeros((nl, n), order="F")
    t = initial_temperature(m["coord_nod2D"], nl, Lx, Ly)
    o.step(1000.0, vel, w, w, t, valuesAB=t.copy(order="F"), hor="UPW1", ver="UPW1", lim="NONE")
    assert fortran_e(t.min(), 14, 6).strip() != golden_lines()[0][1]
