"""Batched form report: 8 uneven periodic meshes x 32 tracers, FP32, one graph launch per step.

    python -m fesom2_tracer_dwarf_amd.batch_report [--steps 100] [--ntr 32]

Meshes from 256x256 to 1024x1024 nodes with 32-64 levels (bottom levels
U{3..nl}), tracers T, S and passive blobs / 0-1 steps / white noise,
MUSCL + QR4C + FCT, analytic velocity at CFL 0.3 (each mesh its own dt).
Prints one JSON line: node-level tracer updates per second over the whole
batch, per-step time, and per-mesh sizes.
"""
from __future__ import annotations

import argparse
import json
import sys
import time

SIZES = [(256, 32), (384, 40), (512, 48), (640, 56), (768, 64), (896, 48), (960, 40), (1024, 64)]


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--ntr", type=int, default=32)
    ap.add_argument("--precision", type=int, default=32)
    args = ap.parse_args()
    from .batch import TracerBatch
    t0 = time.perf_counter()
    b = TracerBatch()
    meshes = []
    for k, (n, nl) in enumerate(SIZES):
        L = 2.0e3 * n
        i, dt = b.add_periodic_mesh(n, n, nl, L, L, seed=10 + k, ntracers=args.ntr, precision=args.precision)
        info = b.info(i)
        meshes.append(dict(nodes=info["nod2D"], nl=nl, active_cells=info["active_cells"], dt=dt))
        print(f"[batch] mesh {k}: {n}x{n}x{nl}, {info['active_cells']} active cells", file=sys.stderr, flush=True)
    setup = time.perf_counter() - t0
    sched = b.autotune(2)  # each mesh keeps its fastest launch schedule (tracers restored)
    b.run(2)  # warm-up (captures the graph)
    ms = b.time(args.steps)
    cells = sum(m["active_cells"] for m in meshes)
    upd = args.ntr * cells * args.steps / (ms * 1e-3)
    print(json.dumps({"workload": "8 periodic meshes 256^2..1024^2, 32-64 levels", "tracers_per_mesh": args.ntr,
                      "precision": args.precision, "steps": args.steps, "ms_per_step": ms / args.steps,
                      "updates_per_s": upd, "active_cells_total": cells, "graph_launches_per_step": 1,
                      "setup_s": setup, "schedules": sched, "meshes": meshes}), flush=True)
    b.close()


if __name__ == "__main__":
    main()
