"""The allocation-free entry point, tracer_do_oce_adv_tra_noalloc, against
the entry point it is a variant of, tracer_do_oce_adv_tra, through the C ABI.

The same step kernels write the same tendencies, so the device arrays the
allocation-free form adds into must equal, bit for bit, the host arrays the
allocating form adds into, from the same starting values; two meshes in a
row, the second smaller (the scratch keeps the first call's contents), and a
mesh above the scratch's size is refused with an error.
"""
import numpy as np
import pytest

from test_gpu_parity import setup_case

pytestmark = pytest.mark.gpu


def _mesh(nx, ny, nl, seed):
    return dict(nx=nx, ny=ny, nl=nl, periodic=True, Lx=5e3 * nx, Ly=5e3 * ny, dz=10.0, stretch=0.5, lev_lo=3,
                lev_hi=nl, seed=seed)


def test_noalloc_equals_allocating_form(api):
    import torch
    print("[gpu] allocation-free tracer_do_oce_adv_tra_noalloc vs tracer_do_oce_adv_tra", flush=True)
    rng = np.random.default_rng(2)
    for kw, schemes in ((_mesh(48, 40, 32, 3), ("MUSCL", "QR4C", "FCT")), (_mesh(20, 16, 12, 5), ("MFCT", "PPM", "NONE"))):
        d, dt, trs = setup_case(api, kw, schemes, ntr=2)
        L, N = d["nl"] - 1, d["nod2D"]
        for tr in (1, 2):
            h0 = np.asfortranarray(rng.normal(size=(L, N)))
            v0 = np.asfortranarray(rng.normal(size=(L, N)))
            ref_h, ref_v = api.do_oce_adv_tra(tr, dt, h0.copy(order="F"), v0.copy(order="F"))
            # (N, L) row-major on the device is the (L, N) column-major layout
            dh = torch.from_numpy(np.ascontiguousarray(h0.T)).cuda()
            dv = torch.from_numpy(np.ascontiguousarray(v0.T)).cuda()
            torch.cuda.synchronize()
            rc = api.lib.tracer_do_oce_adv_tra_noalloc(tr, dt, dh.data_ptr(), dv.data_ptr(), d["nl"], N)
            assert rc == 0, api.lib.tracer_last_error()
            h, v = dh.cpu().numpy().T, dv.cpu().numpy().T
            assert np.array_equal(h, ref_h) and np.array_equal(v, ref_v), f"mesh {kw['nx']}x{kw['ny']} tracer {tr}"
            assert not np.array_equal(h, h0)
            assert np.array_equal(api.get_tracer_values(tr), trs[tr - 1])


def test_noalloc_refuses_a_mesh_above_its_scratch(api):
    import torch
    d, dt, trs = setup_case(api, dict(nx=300, ny=300, nl=64, periodic=True, Lx=1.5e6, Ly=1.5e6, dz=10.0,
                                      stretch=0.0, lev_lo=64, lev_hi=64, seed=0), ("MUSCL", "QR4C", "FCT"), ntr=1)
    api.advect_steps(1, dt)  # the device state exists
    L, N = d["nl"] - 1, d["nod2D"]
    dh = torch.zeros((N, L), dtype=torch.float64, device="cuda")
    dv = torch.zeros((N, L), dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    rc = api.lib.tracer_do_oce_adv_tra_noalloc(1, dt, dh.data_ptr(), dv.data_ptr(), d["nl"], N)
    assert rc == -1 and b"MiB" in api.lib.tracer_last_error()
    assert float(dh.abs().max()) == 0.0
