"""Shared fixtures. GPU tests are marked ``gpu``; everything else runs on CPU."""
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")


MESH_KEYS = None


def mesh_dict(api):
    """All mesh arrays the oracle reads, fetched from the library's setup code."""
    import oracle as orc
    d = api.dims()
    m = {k: api.mesh_array(k) for k in orc.MESH_INT + orc.MESH_DBL}
    m.update(d)
    return m


@pytest.fixture(scope="session")
def api():
    from fesom2_tracer_dwarf_amd import FESOMTracerAdvection
    return FESOMTracerAdvection()


PI_MESH = ROOT / "tests" / "golden" / "pi_mesh"


@pytest.fixture(scope="session")
def pi_mesh_dir(tmp_path_factory):
    """The reference's pi mesh (3140 nodes, 5839 triangles, 48 levels), unpacked."""
    import gzip
    d = tmp_path_factory.mktemp("pi_mesh")
    for f in PI_MESH.glob("*.out.gz"):
        (d / f.name[:-3]).write_bytes(gzip.decompress(f.read_bytes()))
    return d


def write_wheel_mesh(d, spokes=(12, 16), nx=30, ny=24, nl=12, seed=5):
    """FESOM mesh files (nod2d.out, elem2d.out, aux3d.out, elvls.out; edges
    derived by the reader) of a Delaunay triangulation in lon/lat degrees: a
    jittered grid on [0, 7.5] x [-3, 3] with one "wheel" per entry of
    ``spokes`` -- a centre node joined to k nodes on a circle, the grid
    cleared around it -- so the centre has exactly k edges and k elements.
    Element levels are random in [3, nl]; returns the node degrees."""
    from scipy.spatial import Delaunay
    rng = np.random.default_rng(seed)
    h = 7.5 / nx
    gx, gy = np.meshgrid(np.linspace(0.0, 7.5, nx + 1), np.linspace(-3.0, 3.0, ny + 1), indexing="ij")
    pts = np.stack([gx.ravel(), gy.ravel()], 1)
    inner = (pts[:, 0] > 0) & (pts[:, 0] < 7.5) & (np.abs(pts[:, 1]) < 3.0)
    pts[inner] += rng.uniform(-0.2 * h, 0.2 * h, (inner.sum(), 2))
    r0 = 1.1 * h
    centres = [(1.5 + 4.5 * i / max(len(spokes) - 1, 1), 0.3 * (-1) ** i) for i in range(len(spokes))]
    for cx, cy in centres:
        pts = pts[np.hypot(pts[:, 0] - cx, pts[:, 1] - cy) > 1.45 * r0]
    for (cx, cy), k in zip(centres, spokes):
        a = 2 * np.pi * (np.arange(k) + 0.25) / k
        pts = np.vstack([pts, [[cx, cy]], np.stack([cx + r0 * np.cos(a), cy + r0 * np.sin(a)], 1)])
    tri = Delaunay(pts).simplices
    N, T = len(pts), len(tri)
    edges = {tuple(sorted((int(t[i]), int(t[(i + 1) % 3])))) for t in tri for i in range(3)}
    deg = np.zeros(N, int)
    for a, b in edges:
        deg[a] += 1
        deg[b] += 1
    zbar = -np.concatenate([[0.0], np.cumsum(np.linspace(8.0, 20.0, nl - 1))])
    elv = rng.integers(3, nl + 1, T)
    with open(d / "nod2d.out", "w") as f:
        f.write(f"{N}\n")
        for i, (x, y) in enumerate(pts):
            f.write(f"{i + 1} {x:.12f} {y:.12f} 0\n")
    with open(d / "elem2d.out", "w") as f:
        f.write(f"{T}\n")
        for t in tri:
            f.write(f"{t[0] + 1} {t[1] + 1} {t[2] + 1}\n")
    with open(d / "aux3d.out", "w") as f:
        f.write(f"{nl}\n" + "\n".join(f"{z:.6f}" for z in -zbar) + "\n")
        f.write("\n".join(f"{-zbar[-1]:.6f}" for _ in range(N)) + "\n")
    (d / "elvls.out").write_text("\n".join(str(v) for v in elv) + "\n")
    return deg
