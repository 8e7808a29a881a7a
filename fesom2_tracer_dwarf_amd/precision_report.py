"""Precision form report: reduced-precision steps against the FP64 path.

    python -m fesom2_tracer_dwarf_amd.precision_report [--steps 200] [--nx 1000 --ny 1000 --nl 64]

Workload: the periodic 1000x1000x64 mesh (bottom levels U{3..64}, seed 0),
4 tracers (T, S, a 0/1 step function, a Gaussian blob), MUSCL + QR4C + FCT,
analytic velocity at CFL 0.3. For precision 32 (FP32 fluxes, gradients and
limiter; FP64 node tracer, divergence accumulation and bounds) and 16 (the
same with FP16 gradient storage), a second device state in FP64 advances in
lockstep and after every step the wet-layer L-infinity and L2 differences are
reduced on the device. Prints one JSON line per precision with the per-step
relative errors of every tracer and their maxima, plus mass drift and range.
"""
from __future__ import annotations

import argparse
import json
import sys

import numpy as np


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--nx", type=int, default=1000)
    ap.add_argument("--ny", type=int, default=1000)
    ap.add_argument("--nl", type=int, default=64)
    ap.add_argument("--ntr", type=int, default=4)
    args = ap.parse_args()
    from . import FESOMTracerAdvection, configs

    Lx = Ly = 2.0e6 * args.nx / 1000.0
    api = FESOMTracerAdvection()
    api.generate_mesh(args.nx, args.ny, args.nl, periodic=True, Lx=Lx, Ly=Ly, dz=10.0, lev_lo=3,
                      lev_hi=args.nl, seed=0, sfc=True)
    api.allocate_tracers(args.ntr)
    dt = api.make_velocity(0.1, 0.05, 0.2, 0.5, 0.3)
    trs = configs.initial_tracers(api.mesh_array("coord"), args.nl, Lx, Ly, args.ntr, seed=0)
    names = ["T", "S", "step", "blob", "noise"][: args.ntr] + [f"tr{i}" for i in range(5, args.ntr)]
    for prec in (32, 16):
        api.set_option("precision", 64)
        for i, t in enumerate(trs):
            api.set_values(i + 1, t)
            api.set_schemes(i + 1, "MUSCL", "QR4C", "FCT")
        d0 = [api.diagnostics(i + 1) for i in range(args.ntr)]
        api.set_option("precision", prec)
        print(f"[precision] FP{prec} vs FP64, {args.steps} steps", file=sys.stderr, flush=True)
        err = api.precision_error(64, args.steps, dt)
        d1 = [api.diagnostics(i + 1) for i in range(args.ntr)]
        rec = {
            "precision": prec,
            "steps": args.steps,
            "workload": f"periodic_{args.nx}x{args.ny}x{args.nl}",
            "tracers": names,
            "linf_rel_max": err[:, :, 2].max(axis=0).tolist(),
            "l2_rel_max": err[:, :, 3].max(axis=0).tolist(),
            "linf_rel_final": err[-1, :, 2].tolist(),
            "l2_rel_final": err[-1, :, 3].tolist(),
            "linf_rel_per_step": np.round(err[:, :, 2], 12).tolist(),
            "l2_rel_per_step": np.round(err[:, :, 3], 12).tolist(),
            "mass_drift_rel": [abs(b["mass"] - a["mass"]) / abs(a["mass"]) for a, b in zip(d0, d1)],
            "range_before": [(a["active_min"], a["active_max"]) for a in d0],
            "range_after": [(b["active_min"], b["active_max"]) for b in d1],
        }
        print(json.dumps(rec), flush=True)


if __name__ == "__main__":
    main()
