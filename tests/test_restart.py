"""Derived-type binary restarts (tracer_init / tracer_finalize / tracer_load_mesh_partition), host side.

  * the item codec against a real Fortran compiler: tests/fortran/restart_items.f90
    writes files in the t_partit and t_tracer item layout through user-defined
    derived-type output (one parent record, child items); the library parses
    them and writes them back byte for byte (skipped without a Fortran
    compiler);
  * corrupted files (truncated, trailing bytes, a second record) are refused
    with an error;
  * tracer_finalize then tracer_init restores the mesh arrays, the tracers and
    the velocity exactly, and a second tracer_finalize writes the same bytes;
  * tracer_load_mesh_partition reads the mesh and velocity without tracers.

The item order restates the reference's user-defined I/O procedures
(lib/MOD_MESH.F90:185-279, lib/MOD_PARTIT.F90:114-182,
lib/MOD_TRACER.F90:104-239, lib/MOD_DYN.F90:183-290); the reference ships no
restart files, so the order itself is pinned only by these restatements.
"""
import filecmp
import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

HERE = Path(__file__).resolve().parent
FLANG = shutil.which("amdflang") or shutil.which("flang") or (
    "/opt/rocm/llvm/bin/amdflang" if Path("/opt/rocm/llvm/bin/amdflang").exists() else None)
KINDS = ("t_mesh", "t_partit", "t_tracer", "t_dynamics")


def _copy(api, kind, src, dst):
    return api.lib.tracer_restart_file_copy(kind.encode(), str(src).encode(), str(dst).encode())


@pytest.mark.skipif(FLANG is None, reason="no Fortran compiler")
def test_codec_against_fortran_derived_type_output(api, tmp_path):
    exe = tmp_path / "write_items"
    subprocess.run([FLANG, str(HERE / "fortran" / "restart_items.f90"), "-o", str(exe)], check=True,
                   cwd=tmp_path, capture_output=True)
    subprocess.run([str(exe)], check=True, cwd=tmp_path)
    for kind in ("t_partit", "t_tracer"):
        src, dst = tmp_path / f"{kind}.0", tmp_path / f"{kind}.copy"
        raw = src.read_bytes()
        assert int.from_bytes(raw[:4], "little") == len(raw) - 8 == int.from_bytes(raw[-4:], "little")
        assert _copy(api, kind, src, dst) == 0
        assert filecmp.cmp(src, dst, shallow=False), kind


def _restart_dir(api, tmp_path, name):
    d = tmp_path / name
    assert api.finalize(str(d)) == 0
    return d


def _setup(api):
    api.generate_mesh(21, 17, 12, periodic=False, Lx=2.1e5, Ly=1.7e5, dz=10.0, stretch=0.3, lev_lo=3, lev_hi=12,
                      seed=4, sfc=True)
    api.allocate_tracers(3)
    api.make_velocity(0.1, 0.05, 0.2, 0.5, 0.3)
    rng = np.random.default_rng(3)
    d = api.dims()
    vals = [np.asfortranarray(rng.uniform(0, 30, (d["nl"] - 1, d["nod2D"]))) for _ in range(3)]
    for i, (v, sch) in enumerate(zip(vals, [("MUSCL", "QR4C", "FCT"), ("MFCT", "PPM", "FCT"),
                                            ("UPW1", "CDIFF", "NONE")])):
        api.set_values(i + 1, v)
        api.set_schemes(i + 1, *sch, ph=0.9 - 0.1 * i, pv=1.0)
    return d, vals


MESH_KEYS = ["edges", "edge_tri", "elem2D_nodes", "nlevels", "ulevels", "nlevels_nod2D", "ulevels_nod2D",
             "nlevels_nod2D_min", "ulevels_nod2D_max", "nod_in_elem2D_num", "edge_up_dn_tri", "nboundary_lay",
             "zbar", "Z", "depth", "elem_cos", "elem_area", "gradient_sca", "edge_dxdy", "edge_cross_dxdy", "area",
             "areasvol", "hnode", "hnode_new", "helem", "Z_3d_n", "zbar_3d_n", "uv", "w", "w_e"]


def test_finalize_init_round_trip(api, tmp_path):
    d, vals = _setup(api)
    before = {k: api.mesh_array(k) for k in MESH_KEYS}
    nie = api.mesh_array("nod_in_elem2D").reshape(-1, d["maxAdj"])
    d1 = _restart_dir(api, tmp_path, "r1")
    assert sorted(p.name for p in (d1 / "np1").iterdir()) == sorted(f"{k}.0" for k in KINDS)
    assert (d1 / "meta.time").read_text().splitlines()[:3] == ["0", "0", "! year: 0"]
    assert api.initialize(str(d1)) == 0
    assert api.dims() == d
    for k in MESH_KEYS:
        assert np.array_equal(api.mesh_array(k), before[k]), k
    nie2 = api.mesh_array("nod_in_elem2D").reshape(-1, d["maxAdj"])
    num = before["nod_in_elem2D_num"]
    for n in range(d["nod2D"]):
        assert np.array_equal(nie2[n, :num[n]], nie[n, :num[n]])
    for i in range(3):
        assert np.array_equal(api.get_tracer_values(i + 1), vals[i])
    d2 = _restart_dir(api, tmp_path, "r2")
    for k in KINDS:
        assert filecmp.cmp(d1 / "np1" / f"{k}.0", d2 / "np1" / f"{k}.0", shallow=False), k
    for k in KINDS:  # the codec's own output is canonical
        assert _copy(api, k, d1 / "np1" / f"{k}.0", tmp_path / f"{k}.copy") == 0
        assert filecmp.cmp(d1 / "np1" / f"{k}.0", tmp_path / f"{k}.copy", shallow=False), k


def test_load_mesh_partition_and_errors(api, tmp_path):
    d, _ = _setup(api)
    w = api.mesh_array("w")
    d1 = _restart_dir(api, tmp_path, "r")
    assert api.finalize(str(tmp_path / "again")) == -1          # not initialised after tracer_finalize
    assert api.load_mesh_partition(d1) == d
    assert np.array_equal(api.mesh_array("w"), w)
    assert api.initialize(str(tmp_path / "missing")) == -1
    bad = tmp_path / "bad"
    shutil.copytree(d1, bad)
    mesh = bad / "np1" / "t_mesh.0"
    raw = mesh.read_bytes()
    mesh.write_bytes(raw[:len(raw) // 2])
    assert api.initialize(str(bad)) == -1
    assert "truncated" in api_error(api)
    mesh.write_bytes(raw + raw[:12])
    assert api.initialize(str(bad)) == -1
    assert "more than one record" in api_error(api)
    mesh.write_bytes(raw)
    part = bad / "np1" / "t_partit.0"
    payload = part.read_bytes()[4:-4] + b"\0\0\0\0"           # one item too many inside the record
    marker = len(payload).to_bytes(4, "little")
    part.write_bytes(marker + payload + marker)
    assert api.initialize(str(bad)) == -1
    assert "left after the last item" in api_error(api)


def api_error(api):
    from fesom2_tracer_dwarf_amd import _lib
    return _lib.last_error(api.lib)


def test_round_trip_on_the_pi_mesh(api, pi_mesh_dir, tmp_path):
    """The reference's pi mesh (coasts, so boundary edges with edge_tri(2) = 0 in the file; variable bottom)."""
    api.set_option("force_rotation", 1)
    d = api.load_mesh_from_files(pi_mesh_dir)
    api.allocate_tracers(2)
    api.make_velocity(0.1, 0.0, 0.2, 0.5, 0.3)
    rng = np.random.default_rng(8)
    vals = [np.asfortranarray(rng.uniform(-2, 25, (d["nl"] - 1, d["nod2D"]))) for _ in range(2)]
    for i, v in enumerate(vals):
        api.set_values(i + 1, v)
        api.set_schemes(i + 1, "MUSCL", "PPM", "FCT")
    before = {k: api.mesh_array(k) for k in MESH_KEYS}
    assert np.any(before["edge_tri"].reshape(-1, 2)[:, 1] == -1)
    r = _restart_dir(api, tmp_path, "pi")
    raw = (r / "np1" / "t_mesh.0").read_bytes()
    assert int.from_bytes(raw[:4], "little", signed=True) == len(raw) - 8
    assert api.initialize(str(r)) == 0
    assert api.dims() == d
    for k in MESH_KEYS:
        assert np.array_equal(api.mesh_array(k), before[k]), k
    for i in range(2):
        assert np.array_equal(api.get_tracer_values(i + 1), vals[i])
