"""Order of generated meshes (tracer_generate_mesh with sfc != 0): nodes and
elements follow a Hilbert curve over the grid.

  * on a 2^k x 2^k grid, consecutive nodes are grid neighbours, one step in x
    or in y; a Morton (Z) order breaks this at every quadrant boundary (the
    check is run on a Morton order of the same grid and fails there);
  * the elements come in cell pairs (the cell's two triangles, both listed
    from the cell's lower-left node), and consecutive cells are grid
    neighbours on that grid;
  * on any rectangle the order is a permutation of the grid points.
"""
import numpy as np
import pytest


def _grid_index(api, nx, ny, periodic):
    d = api.generate_mesh(nx, ny, 4, periodic=periodic, Lx=float(nx), Ly=float(ny), dz=1.0, lev_lo=4, lev_hi=4,
                          seed=0, sfc=True)
    xy = api.mesh_array("coord_nod2D").reshape(-1, 2)
    step = xy[xy > 0].min()  # one grid step (radians on the planar mesh's scale)
    ij = np.rint(xy / step).astype(np.int64)
    return d, ij


def _morton_order(n):
    def spread(v):
        out = 0
        for b in range(16):
            out |= ((v >> b) & 1) << (2 * b)
        return out
    pts = [(i, j) for j in range(n) for i in range(n)]
    return np.array(sorted(pts, key=lambda p: spread(p[0]) | (spread(p[1]) << 1)))


def _steps_are_unit(ij):
    return np.all(np.abs(np.diff(ij, axis=0)).sum(axis=1) == 1)


@pytest.mark.parametrize("n", [8, 16, 32])
def test_consecutive_nodes_are_grid_neighbours(api, n):
    d, ij = _grid_index(api, n, n, periodic=True)
    assert d["nod2D"] == n * n
    assert _steps_are_unit(ij)
    assert not _steps_are_unit(_morton_order(n))  # the check tells the two curves apart


def test_consecutive_cells_are_grid_neighbours(api):
    n = 16
    d, ij = _grid_index(api, n, n, periodic=True)
    el = api.mesh_array("elem2D_nodes").reshape(-1, 3)
    assert el.shape[0] == 2 * n * n
    assert np.array_equal(el[0::2, 0], el[1::2, 0])  # a cell's two triangles start at its lower-left node
    assert _steps_are_unit(ij[el[0::2, 0]])


@pytest.mark.parametrize("nx,ny,periodic", [(12, 10, True), (13, 7, False), (40, 3, True)])
def test_order_is_a_permutation_of_the_grid(api, nx, ny, periodic):
    d, ij = _grid_index(api, nx, ny, periodic)
    assert d["nod2D"] == nx * ny
    flat = ij[:, 1] * nx + ij[:, 0]
    assert np.array_equal(np.sort(flat), np.arange(nx * ny))
