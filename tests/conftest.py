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
