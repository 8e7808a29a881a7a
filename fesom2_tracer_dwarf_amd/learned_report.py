"""Learned-limiter report: train on the FCT's own pairs, then run it at scale.

    python -m fesom2_tracer_dwarf_amd.learned_report [--pairs 50000000] [--steps 200]

1. Training set: a periodic 512x512x32 mesh (bottom levels U{3..32}) with 4
   passive tracers of seeded random fields (white noise, 0/1 steps, Gaussian
   blobs), MUSCL + QR4C + FCT in FP32; every FCT step appends one row per
   edge level (10 features, Zalesak's coefficient) until --pairs rows.
2. Training: Adam on the device (BF16 matrix-core forward/backward, FP32
   master weights), --epochs epochs of --batch rows.
3. Inference: the periodic 1000x1000x64 mesh, 4 tracers (T, S, a 0/1 step
   function, a blob), FP32, --steps steps with the learned limiter and, for
   reference, with Zalesak (FP32 and FP64). Reports the distance of the
   learned result to Zalesak's, the fallback fraction, mass drift, range, and
   the per-step time of the learned FP32 step against the FP64 Zalesak step.
Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np


def log(msg):
    print(f"[learned] {msg}", file=sys.stderr, flush=True)


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--pairs", type=int, default=50_000_000)
    ap.add_argument("--epochs", type=int, default=8)
    ap.add_argument("--batch", type=int, default=262144)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--nx", type=int, default=1000)
    ap.add_argument("--nl", type=int, default=64)
    args = ap.parse_args()
    from . import FESOMTracerAdvection, configs

    api = FESOMTracerAdvection()
    # ---- training pairs on random fields
    t0 = time.perf_counter()
    api.generate_mesh(512, 512, 32, periodic=True, Lx=1.0e6, Ly=1.0e6, dz=10.0, lev_lo=3, lev_hi=32, seed=5, sfc=True)
    api.allocate_tracers(4)
    dt_tr = api.make_velocity(0.1, 0.05, 0.2, 0.5, 0.3)
    coord = api.mesh_array("coord")
    rng = np.random.default_rng(123)
    fields = [configs.tracer_noise(coord, 31, rng), configs.tracer_step(coord, 31, 1.0e6, 1.0e6, rng),
              configs.tracer_blob(coord, 31, 1.0e6, 1.0e6, rng, depth_decay=False), configs.tracer_noise(coord, 31, rng)]
    for i, f in enumerate(fields):
        api.set_values(i + 1, f)
        api.set_schemes(i + 1, "MUSCL", "QR4C", "FCT")
    api.set_option("precision", 32)
    api.learn_clear()
    rows, cap = 0, int(args.pairs * 1.5)
    while rows < args.pairs:
        before = rows
        rows = api.learn_collect(1, dt_tr, cap)
        if rows == before:
            break
    log(f"{rows} training rows in {time.perf_counter() - t0:.1f} s")
    t1 = time.perf_counter()
    losses = api.learn_train(args.epochs, batch=args.batch, lr=args.lr, seed=0)
    t_train = time.perf_counter() - t1
    log(f"trained {args.epochs} epochs in {t_train:.1f} s: loss {np.round(losses, 5).tolist()}")
    params = api.mlp_params().copy()

    # ---- inference at scale
    Lx = 2.0e6 * args.nx / 1000.0
    api.generate_mesh(args.nx, args.nx, args.nl, periodic=True, Lx=Lx, Ly=Lx, dz=10.0, lev_lo=3, lev_hi=args.nl,
                      seed=0, sfc=True)
    api.allocate_tracers(4)
    dt = api.make_velocity(0.1, 0.05, 0.2, 0.5, 0.3)
    trs = configs.initial_tracers(api.mesh_array("coord"), args.nl, Lx, Lx, 4, seed=0)
    api.set_mlp_params(params)

    def run(prec, lim):
        api.set_option("precision", prec)
        for i, t in enumerate(trs):
            api.set_values(i + 1, t)
            api.set_schemes(i + 1, "MUSCL", "QR4C", lim)
        d0 = [api.diagnostics(i + 1) for i in range(4)]
        api.advect_steps(args.steps, dt)
        d1 = [api.diagnostics(i + 1) for i in range(4)]
        vals = [api.get_tracer_values(i + 1) for i in range(4)]
        fb = [api.learn_fallback(i + 1) for i in range(4)] if lim == "LEARNED" else None
        ms = api.time_steps(10, dt) / 10.0  # timing on the evolved state (does not feed the comparison)
        return vals, d0, d1, fb, ms

    log("Zalesak FP64")
    z64, _, _, _, ms_z64 = run(64, "FCT")
    log("Zalesak FP32")
    z32, _, _, _, ms_z32 = run(32, "FCT")
    log("learned FP32")
    ln, d0, d1, fb, ms_ln = run(32, "LEARNED")
    names = ["T", "S", "step", "blob"]
    rel = lambda a, b: float(np.linalg.norm(a - b) / np.linalg.norm(b))  # noqa: E731
    rec = {
        "training": {"rows": rows, "mesh": "periodic_512x512x32", "epochs": args.epochs, "batch": args.batch,
                     "lr": args.lr, "loss_per_epoch": losses.tolist(), "train_s": t_train},
        "workload": f"periodic_{args.nx}x{args.nx}x{args.nl}", "steps": args.steps, "tracers": names,
        "l2_rel_learned_vs_zalesak_fp32": [rel(a, b) for a, b in zip(ln, z32)],
        "l2_rel_learned_vs_zalesak_fp64": [rel(a, b) for a, b in zip(ln, z64)],
        "l2_rel_zalesak_fp32_vs_fp64": [rel(a, b) for a, b in zip(z32, z64)],
        "fallback_fraction_last_step": fb,
        "mass_drift_rel": [abs(b["mass"] - a["mass"]) / abs(a["mass"]) for a, b in zip(d0, d1)],
        "range_before": [(a["active_min"], a["active_max"]) for a in d0],
        "range_after": [(b["active_min"], b["active_max"]) for b in d1],
        "ms_per_step": {"zalesak_fp64": ms_z64, "zalesak_fp32": ms_z32, "learned_fp32": ms_ln},
        "learned_fp32_speedup_over_fp64": ms_z64 / ms_ln,
    }
    print(json.dumps(rec), flush=True)


if __name__ == "__main__":
    main()
