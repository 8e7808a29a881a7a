#!/usr/bin/env python3
"""Benchmark of the tracer-advection step on one MI355X.

Workload (default): doubly periodic 1000 x 1000 node mesh (3.0M edges, 2.0M
triangles), 64 level interfaces with per-node bottom level drawn from
U{3..64} (seed 0), tracers T and S, MUSCL + QR4C + FCT (Zalesak), analytic
divergence-free velocity at CFL 0.3. One "step" advances both tracers once.

Metric: node-level tracer updates per second = n_tracers * C * steps / t,
C = active (node, layer) cells of the generated mesh. The timed region is K
graph-replayed steps bracketed by hipEvents on the library's stream (inputs
already resident in HBM). FP64 is the headline (the parity path); the FP32
form (FP64 node tracer/accumulation) is reported beside it, and so is the
rate of a caller that hands the tracers over as host arrays every step
("host_round_trip": upload, one step, download over PCIe).

Roofline: the step's compulsory HBM bytes (every semantic input read once and
every output written once, at its stored precision) over the measured step
time, against the 8 TB/s peak; the measured step traffic (rocprofv3
FETCH_SIZE + WRITE_SIZE, copy-calibrated), the busiest kernel's own figure and
the rate of the library's hand-written streaming copy (the practical ceiling)
are reported beside it. CPU baseline: the oracle (a C restatement of the
reference's FP64 path) on the host cores -- all cores through its OpenMP build,
one thread, and the 20x20x10 analytic case -- on bounded samples. Beside the
headline the line also carries the batched form (8 meshes x 32 tracers, one
graph launch per step) and the learned limiter's step time against Zalesak's.

Prints ONE JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "node-level tracer updates/s at 1M nodes×64 lvls, FP64 vs FP32; % HBM peak"
HBM_PEAK_GBS = 8000.0
# HBM bytes per dispatch measured with rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE on
# this workload (copy-calibrated), for the roofline's "traffic" field
TRAFFIC_TABLE = ROOT / "profiles" / "hbm_traffic_periodic_1000x1000x64.md"


def counts(api):
    m = {k: api.mesh_array(k) for k in ("nlevels_nod2D", "nlevels", "edge_tri")}
    d = api.dims()
    C = int((m["nlevels_nod2D"] - 1).sum())
    Ce = int((m["nlevels"] - 1).sum())
    et = m["edge_tri"].reshape(-1, 2)
    nl1 = m["nlevels"][et[:, 0]] - 1
    nl2 = np.where(et[:, 1] >= 0, m["nlevels"][np.maximum(et[:, 1], 0)] - 1, 0)
    Ced = int(np.maximum(nl1, nl2).sum())
    return dict(C=C, Ce=Ce, Ced=Ced, N=d["nod2D"], T=d["elem2D"], E=d["edge2D"], nl=d["nl"], maxAdj=d["maxAdj"])


def step_alg_bytes(c, ntr, s_t, s_v, s_m):
    """Compulsory HBM bytes of one step: each semantic input read once, each output written once."""
    return (ntr * 2 * s_t * c["C"]                       # tracer read + write
            + s_v * 2 * c["Ce"] + s_m * c["Ce"]           # uv(2,.,elem) + helem
            + s_v * (c["C"] + c["N"])                     # w on interfaces
            + s_m * 5 * c["C"]                            # hnode, hnode_new, areasvol, Z_3d_n, zbar_3d_n
            + c["E"] * (16 + 6 * s_m) + c["T"] * (20 + 7 * s_m) + 8 * c["N"] + 4 * c["maxAdj"] * c["N"])


def kernel_alg_bytes(name, c, ntr, s_t, s_r, s_g, we_alias=True, h_alias=True):
    """Compulsory bytes of one step's launch of a kernel (all ntr tracers
    batched into it): every array it reads or writes counted once over its
    active cells (C node layers, C+N node interfaces, Ce element layers, Ced
    edge levels); mesh and velocity terms once, tracer terms ntr times.
    Gradients count the nodal means only (the up/down-wind triangle values
    replace them on shared levels), so the figure is a lower bound. w_e shares
    w's storage and hnode_new hnode's when they are equal (we_alias, h_alias),
    as on the benchmark mesh; the FP64 copies of areasvol, hnode, hnode_new
    that the FP32 form keeps for its FP64 accumulation are counted as read."""
    C, Ce, Ced, N, T, E = c["C"], c["Ce"], c["Ced"], c["N"], c["T"], c["E"]
    CI = C + N
    h64 = 8 * C if h_alias else 16 * C          # FP64 hnode (and hnode_new)
    if name == "k_gnod":          # valuesAB -> nodal mean gradients (2 comps); 5 fan weights per element around
                                  # each node (3T in all), 1-ring records
        return ntr * (s_t * C + 2 * s_g * C) + 5 * s_r * 3 * T + 48 * N
    if name == "k_edge":          # tracer (T and AB share storage), nodal gradients, vel, helem, edge records and
                                  # up/down-wind weights -> transport, HO-LO
        return ntr * (s_t * C + 2 * s_g * C + s_r * Ced) + s_r * Ced + 3 * s_r * Ce + E * (48 + 16 * s_r)
    if name == "k_node_lo":
        # edge transport, tracer -> LO (FP64), adf_v; on interfaces w (w_e shares its storage when equal),
        # area, zbar_3d_n; on layers Z_3d_n and the FP64 hnode (hnode_new), and the FP64 areasvol copy in FP32
        # (FP64 reads area, equal to it on wet layers); node->edge records
        n_if = 3 if we_alias else 4
        geo = s_r * C + h64 + (8 * C if s_r == 4 else 0)
        return ntr * (s_t * C + 8 * C + s_r * CI) + s_r * Ced + n_if * s_r * CI + geo + 80 * N
    if name == "k_node_limit":    # HO-LO edge fluxes, adf_v, LO, tracer -> R+/R-; areasvol, 1-ring and
                                  # node->edge records
        return ntr * (s_r * Ced + s_r * CI + 8 * C + s_t * C + 2 * s_r * C) + s_r * C + 128 * N
    if name == "k_node_update":   # HO-LO edge flux, adf_v, R+/R- (own and 1-ring), LO, tracer -> tracer;
                                  # FP64 areasvol, hnode (hnode_new)
        return ntr * (s_r * Ced + s_r * CI + 2 * s_r * C + 8 * C + 2 * s_t * C) + 8 * C + h64 + 80 * N
    if name == "k_node_nolim":
        return ntr * (s_r * Ced + 2 * s_t * C) + 2 * s_r * CI + 2 * s_r * C + 48 * N
    return 0


def run_precision(api, prec, dt, steps, warmup, ntr, tune=True):
    api.set_option("precision", prec)
    table = api.autotune(dt, reps=5) if tune else []
    api.advect_steps(warmup, dt)
    t0 = time.perf_counter()
    ms = api.time_steps(steps, dt)
    wall = time.perf_counter() - t0
    kern = api.time_kernels(dt, reps=3)
    return ms, wall, kern, table


def host_round_trip(api, dt, trs, c, reps=3):
    """The rate a caller sees when it hands the tracers over as host arrays every
    step: tracer_set_values of every tracer, one step, tracer_get_values of every
    tracer (pageable numpy arrays, FP64), timed on the wall clock. Reported beside
    the headline, never as it: the headline keeps the inputs resident in HBM."""
    ntr = len(trs)
    host_bytes = 2 * sum(t.nbytes for t in trs)
    out = None
    t_total = 0.0
    for r in range(reps + 1):
        t0 = time.perf_counter()
        for i, t in enumerate(trs):
            api.set_values(i + 1, t)
        api.advect_steps(1, dt)
        out = [api.get_tracer_values(i + 1) for i in range(ntr)]
        if r > 0:  # the first pass warms the host pages
            t_total += time.perf_counter() - t0
    ms = t_total / reps * 1e3
    assert all(np.all(np.isfinite(o)) for o in out)
    return dict(updates_per_s=ntr * c["C"] / (ms * 1e-3), ms_per_step=ms, host_bytes_per_step=host_bytes,
                host_gbs=host_bytes / (ms * 1e-3) / 1e9, steps=reps,
                what="set_values + 1 step + get_values of every tracer, FP64, pageable host arrays")


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or platform.machine()


def host_threads():
    """Threads for the all-cores leg: the CPU share this process may use."""
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit() and int(env) > 0:
        n = min(n, int(env))
    return max(1, n)


def oracle_leg(api, dt, trs, threads, max_seconds, max_steps, hor="MUSCL", ver="QR4C", lim="FCT"):
    """Time single-tracer oracle steps of the api's current mesh (tracers in
    turn) until max_seconds or max_steps; returns (steps, seconds)."""
    import oracle as orc
    mesh = {k: api.mesh_array(k) for k in orc.MESH_INT + orc.MESH_DBL}
    mesh.update(api.dims())
    o = orc.Oracle(mesh, threads=threads)
    uv, w, we = api.mesh_array("uv"), api.mesh_array("w"), api.mesh_array("w_e")
    vals = [np.array(t, order="F", copy=True) for t in trs]
    done, t_total = 0, 0.0
    while True:
        t0 = time.perf_counter()
        o.step(dt, uv, w, we, vals[done % len(vals)], hor=hor, ver=ver, lim=lim)
        t_total += time.perf_counter() - t0
        done += 1
        if t_total >= max_seconds or done >= max_steps:
            break
    return done, t_total, o.threads


def cpu_baseline(api, dt, ntr, c, max_seconds=20.0):
    """The oracle (the CPU FP64 restatement of the reference path) on the host
    cores, on bounded samples: config 1 on all cores (OpenMP build, bitwise the
    serial result) and on one thread, and config 0 (the analytic dwarf, 20x20
    periodic nodes x 10 levels, 2 tracers, 10 steps) on one thread. Leaves the
    library holding the config-0 mesh."""
    from fesom2_tracer_dwarf_amd import configs
    cfg = configs.WORKLOADS["periodic_1000x1000x64"]
    trs = configs.initial_tracers(api.mesh_array("coord"), cfg["nl"], cfg["Lx"], cfg["Ly"], ntr, seed=0)
    model = cpu_model()
    nthr = host_threads()
    log(f"cpu baseline: all cores ({nthr} threads)")
    n_all, t_all, used = oracle_leg(api, dt, trs, nthr, 0.6 * max_seconds, 16)
    log(f"cpu all cores: {n_all} tracer-steps in {t_all:.1f} s")
    log("cpu baseline: 1 thread")
    n_one, t_one, _ = oracle_leg(api, dt, trs, None, 0.4 * max_seconds, 4)
    log(f"cpu 1 thread: {n_one} tracer-steps in {t_one:.1f} s")
    # config 0 on one thread, whole (10 steps of both tracers)
    c0 = configs.WORKLOADS["analytic_20x20x10"]
    api.generate_mesh(c0["nx"], c0["ny"], c0["nl"], periodic=True, Lx=c0["Lx"], Ly=c0["Ly"], dz=c0["dz"],
                      lev_lo=c0["lev"][0], lev_hi=c0["lev"][1], seed=0, sfc=False)
    api.allocate_tracers(c0["ntr"])
    dt0 = api.make_velocity(0.1, 0.05, 0.2, 0.5, 0.3)
    trs0 = configs.initial_tracers(api.mesh_array("coord"), c0["nl"], c0["Lx"], c0["Ly"], c0["ntr"], seed=0)
    C0 = int((api.mesh_array("nlevels_nod2D") - 1).sum())
    n0, t0, _ = oracle_leg(api, dt0, trs0, None, 1e9, c0["steps"] * c0["ntr"])
    return dict(value=n_all * c["C"] / t_all, unit="updates/s", cores=used, kind="port",
                cpu=model, nproc=os.cpu_count(),
                sample=f"{n_all} single-tracer MUSCL+QR4C+FCT steps of periodic_1000x1000x64 ({t_all:.1f} s) on "
                       f"{used} threads of the OpenMP oracle build (deterministic, equal to the serial build bit for "
                       f"bit)",
                one_thread=dict(value=n_one * c["C"] / t_one, unit="updates/s", cores=1,
                                sample=f"{n_one} single-tracer steps of periodic_1000x1000x64 ({t_one:.1f} s), "
                                       "serial oracle"),
                config0=dict(value=n0 * C0 / t0, unit="updates/s", cores=1, ms_per_step=t0 / c0["steps"] * 1e3,
                             sample=f"analytic_20x20x10: {c0['steps']} steps x {c0['ntr']} tracers, MUSCL+QR4C+FCT, "
                                    f"serial oracle ({t0 * 1e3:.1f} ms)"))


BATCH_SIZES = [(256, 32), (384, 40), (512, 48), (640, 56), (768, 64), (896, 48), (960, 40), (1024, 64)]


def batched_form(steps):
    """The batched form: 8 periodic meshes of 256^2 .. 1024^2 nodes with 32..64
    levels, 32 tracers each (T, S, blobs, 0/1 steps, white noise), FP32,
    MUSCL+QR4C+FCT, one graph launch per step for all meshes."""
    from fesom2_tracer_dwarf_amd.batch import TracerBatch
    t0 = time.perf_counter()
    b = TracerBatch()
    cells = 0
    for k, (n, nl) in enumerate(BATCH_SIZES):
        i, _ = b.add_periodic_mesh(n, n, nl, 2.0e3 * n, 2.0e3 * n, seed=10 + k, ntracers=32, precision=32)
        cells += b.info(i)["active_cells"]
    setup = time.perf_counter() - t0
    sched = b.autotune(2)  # each mesh keeps its fastest launch schedule (tracers restored)
    b.run(2)  # warm-up (captures the graph)
    ms = b.time(steps)
    b.close()
    return dict(workload="8 periodic meshes 256^2..1024^2 nodes, 32..64 levels, 32 tracers each, FP32",
                steps=steps, ms_per_step=ms / steps, updates_per_s=32 * cells * steps / (ms * 1e-3),
                active_cells=cells, graph_launches_per_step=1, setup_s=setup,
                schedules=[dict(block_waves=w, batch_tracers=nt) for w, nt in sched])


def learned_form(api, steps):
    """The learned limiter on the 1000x1000x64 mesh with 4 tracers (T, S, a
    0/1 step, a blob), FP32, after training on the FCT's own pairs of a
    256x256x32 mesh of random fields; timed beside Zalesak's FCT in FP64 and
    FP32 from the same state, and its distance to Zalesak FP32 after the
    same number of steps."""
    from fesom2_tracer_dwarf_amd import configs
    t0 = time.perf_counter()
    api.generate_mesh(256, 256, 32, periodic=True, Lx=5.12e5, Ly=5.12e5, dz=10.0, lev_lo=3, lev_hi=32, seed=5,
                      sfc=True)
    api.set_option("precision", 32)
    api.allocate_tracers(4)
    dtt = api.make_velocity(0.1, 0.05, 0.2, 0.5, 0.3)
    coord = api.mesh_array("coord")
    rng = np.random.default_rng(123)
    fields = [configs.tracer_noise(coord, 31, rng), configs.tracer_step(coord, 31, 5.12e5, 5.12e5, rng),
              configs.tracer_blob(coord, 31, 5.12e5, 5.12e5, rng, depth_decay=False),
              configs.tracer_noise(coord, 31, rng)]
    for i, f in enumerate(fields):
        api.set_values(i + 1, f)
        api.set_schemes(i + 1, "MUSCL", "QR4C", "FCT")
    api.learn_clear()
    rows = api.learn_collect(2, dtt, 40_000_000)
    losses = api.learn_train(4, batch=65536, lr=2e-3, seed=0)
    params = api.mlp_params().copy()
    t_train = time.perf_counter() - t0
    cfg = configs.WORKLOADS["periodic_1000x1000x64"]
    api.generate_mesh(cfg["nx"], cfg["ny"], cfg["nl"], periodic=True, Lx=cfg["Lx"], Ly=cfg["Ly"], dz=cfg["dz"],
                      lev_lo=cfg["lev"][0], lev_hi=cfg["lev"][1], seed=0, sfc=True)
    api.allocate_tracers(4)
    dt = api.make_velocity(0.1, 0.05, 0.2, 0.5, 0.3)
    trs = configs.initial_tracers(api.mesh_array("coord"), cfg["nl"], cfg["Lx"], cfg["Ly"], 4, seed=0)
    api.set_mlp_params(params)
    C = int((api.mesh_array("nlevels_nod2D") - 1).sum())

    def run(prec, lim):
        api.set_option("precision", prec)
        for i, t in enumerate(trs):
            api.set_values(i + 1, t)
            api.set_schemes(i + 1, "MUSCL", "QR4C", lim)
        api.advect_steps(steps, dt)
        vals = [api.get_tracer_values(i + 1) for i in range(4)]
        fb = [api.learn_fallback(i + 1) for i in range(4)] if lim == "LEARNED" else None
        ms = api.time_steps(steps, dt) / steps
        return vals, fb, ms

    _, _, ms64 = run(64, "FCT")
    z32, _, ms32 = run(32, "FCT")
    ln, fb, msl = run(32, "LEARNED")
    api.set_option("precision", 64)
    return dict(workload="periodic_1000x1000x64, 4 tracers (T, S, 0/1 step, blob), FP32", steps=steps,
                training=dict(mesh="periodic_256x256x32 random fields", rows=int(rows), epochs=4,
                              loss_first_last=[float(losses[0]), float(losses[-1])], seconds=t_train),
                ms_per_step=dict(zalesak_fp64=ms64, zalesak_fp32=ms32, learned_fp32=msl),
                updates_per_s=dict(zalesak_fp64=4 * C / (ms64 * 1e-3), zalesak_fp32=4 * C / (ms32 * 1e-3),
                                   learned_fp32=4 * C / (msl * 1e-3)),
                learned_speed_over_fp64=ms64 / msl, learned_speed_over_zalesak_fp32=ms32 / msl,
                l2_rel_to_zalesak_fp32=[float(np.linalg.norm(a - b) / np.linalg.norm(b)) for a, b in zip(ln, z32)],
                fallback_fraction=fb)


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--nx", type=int, default=1000)
    ap.add_argument("--ny", type=int, default=1000)
    ap.add_argument("--nl", type=int, default=64)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--no-tune", action="store_true", help="keep the default launch schedule (no autotuning)")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the host round trip, the copy ceiling, the batched form and the learned limiter "
                         "(profiled runs of the step kernels)")
    ap.add_argument("--extra-steps", type=int, default=10,
                    help="timed steps of the batched form and of the learned-limiter comparison")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a plain 1 GiB device-to-device copy (HBM counter calibration)")
    ap.add_argument("--copy-bandwidth", action="store_true",
                    help="only time the library's streaming copy of 4 GiB (the practical HBM ceiling)")
    args = ap.parse_args()
    if args.copy_bandwidth:
        from fesom2_tracer_dwarf_amd import FESOMTracerAdvection
        print(json.dumps(FESOMTracerAdvection().hbm_copy_rate(4 << 30, 20)))
        return
    if args.calibrate_copy:
        import torch
        a = torch.ones(1 << 27, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_each_way": a.numel() * 8, "launches": 5}))
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: the library has no multi-GPU path (the mesh would shard by "
              "space-filling-curve ranges with halo exchanges; see the design notes).", file=sys.stderr)
        sys.exit(2)

    from fesom2_tracer_dwarf_amd import FESOMTracerAdvection, configs
    api = FESOMTracerAdvection()
    cfg = dict(configs.WORKLOADS["periodic_1000x1000x64"])
    cfg["nx"], cfg["ny"], cfg["nl"] = args.nx, args.ny, args.nl
    ntr = cfg["ntr"]
    t_setup = time.perf_counter()
    api.set_option("device_geometry", 1)  # planar mesh: the same arrays as the host derivation, bit for bit
    api.generate_mesh(cfg["nx"], cfg["ny"], cfg["nl"], periodic=True, Lx=cfg["Lx"], Ly=cfg["Ly"], dz=cfg["dz"],
                      lev_lo=cfg["lev"][0], lev_hi=cfg["lev"][1], seed=0, sfc=True)
    api.allocate_tracers(ntr)
    dt = api.make_velocity(0.1, 0.05, 0.2, 0.5, 0.3)
    trs = configs.initial_tracers(api.mesh_array("coord"), cfg["nl"], cfg["Lx"], cfg["Ly"], ntr, seed=0)
    for i, t in enumerate(trs):
        api.set_values(i + 1, t)
        api.set_schemes(i + 1, "MUSCL", "QR4C", "FCT")
    c = counts(api)
    t_setup = time.perf_counter() - t_setup
    log(f"mesh ready in {t_setup:.1f} s: {c}")

    res = {}
    for prec, tag in ((64, "fp64"), (32, "fp32")):
        log(f"timing {tag}")
        ms, wall, kern, table = run_precision(api, prec, dt, args.steps, args.warmup, ntr, not args.no_tune)
        best = min(table, key=lambda r: r["ms_per_step"]) if table else None
        log(f"{tag}: {ms / args.steps:.3f} ms/step, schedule {best}, kernels {kern}")
        s_r = 8 if prec == 64 else 4
        upd = ntr * c["C"] * args.steps / (ms * 1e-3)
        balg = step_alg_bytes(c, ntr, 8, s_r, s_r)
        kb = {k: kernel_alg_bytes(k, c, ntr, 8, s_r, s_r) for k in kern}
        dom = max(kern, key=kern.get)
        res[tag] = dict(updates_per_s=upd, ms_per_step=ms / args.steps, wall_s=wall,
                        step_alg_bytes=balg, step_gbs=balg / (ms / args.steps * 1e-3) / 1e9,
                        step_hbm_frac=balg / (ms / args.steps * 1e-3) / 1e9 / HBM_PEAK_GBS,
                        kernels_ms=kern,
                        kernels_gbs={k: kb[k] / (kern[k] * 1e-3) / 1e9 for k in kern if kern[k] > 0},
                        dominant=dom,
                        schedule=best, autotune=table)
        diag = api.diagnostics(1)
        res[tag]["T_mass"] = diag["mass"]
        res[tag]["nan"] = diag["nan_count"]

    api.set_option("precision", 64)
    host_rt = None
    if not args.no_extras:
        log("host round trip (tracers handed over as host arrays every step)")
        host_rt = host_round_trip(api, dt, trs, c)
        log(f"host round trip: {host_rt['ms_per_step']:.1f} ms/step")

    f64 = res["fp64"]
    dom = f64["dominant"]
    kern_traffic = step_traffic = None
    if (cfg["nx"], cfg["ny"], cfg["nl"]) == (1000, 1000, 64):
        from fesom2_tracer_dwarf_amd.rocprof_summary import read_traffic
        tt = read_traffic(TRAFFIC_TABLE)
        kern_traffic = tt.get(dom, {}).get("fp64")
        # the gradient kernel runs as k_gnod_fan on meshes whose nodes all have an element fan
        look = lambda k: tt.get(k, {}).get("fp64", tt.get(k + "_fan", {}).get("fp64"))  # noqa: E731
        kern_traffic = look(dom)
        per = [look(k) for k in f64["kernels_ms"]]
        step_traffic = sum(per) if per and all(v is not None for v in per) else None
    ceiling = dict(gbs=float("nan"))
    if not args.no_extras:
        log("copy ceiling (hand-written streaming copy)")
        ceiling = api.hbm_copy_rate(4 << 30, 10)
        log(f"copy ceiling: {ceiling}")
    line = {
        "metric": METRIC,
        "value": f64["updates_per_s"],
        "unit": "updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": f64["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"periodic_{cfg['nx']}x{cfg['ny']}x{cfg['nl']}", "nodes": c["N"],
                   "edges": c["E"], "elements": c["T"], "level_interfaces": cfg["nl"],
                   "bottom_levels": "U{3..64} seed 0", "tracers": ntr, "schemes": "MUSCL+QR4C+FCT",
                   "active_node_cells": c["C"], "dense_node_cells": c["N"] * (cfg["nl"] - 1),
                   "dt_s": dt, "cfl": 0.3},
        # the metric's roofline: the step's compulsory bytes (each semantic
        # input read once, each output written once) over the step time
        "roofline": {"bound": "hbm", "achieved": f64["step_gbs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": f64["step_gbs"] / HBM_PEAK_GBS, "traffic": step_traffic,
                     "alg_bytes": f64["step_alg_bytes"], "scope": "whole step (all kernels, both tracers)",
                     "traffic_source": str(TRAFFIC_TABLE.relative_to(ROOT)) if step_traffic else None,
                     "copy_ceiling_gbs": ceiling["gbs"], "frac_of_copy_ceiling": f64["step_gbs"] / ceiling["gbs"]},
        "kernel_roofline": {"kernel": dom, "achieved": f64["kernels_gbs"][dom], "peak": HBM_PEAK_GBS,
                            "unit": "GB/s", "frac": f64["kernels_gbs"][dom] / HBM_PEAK_GBS,
                            "alg_bytes": kernel_alg_bytes(dom, c, ntr, 8, 8, 8), "traffic": kern_traffic,
                            "ms": f64["kernels_ms"][dom]},
        "measured_traffic_gbs": (step_traffic / (f64["ms_per_step"] * 1e-3) / 1e9) if step_traffic else None,
        "copy_ceiling": ceiling,
        "variants": {k: {kk: vv for kk, vv in v.items() if kk not in ("wall_s",)} for k, v in res.items()},
        "fp32_over_fp64": res["fp32"]["updates_per_s"] / f64["updates_per_s"],
        "host_round_trip": host_rt,
        "setup_s": t_setup,
    }
    if not args.no_cpu:
        log("cpu baseline (oracle)")
        api.set_option("precision", 64)
        line["cpu_baseline"] = cpu_baseline(api, dt, ntr, c, args.cpu_seconds)
    else:
        line["cpu_baseline"] = None
    if not args.no_extras:
        log("batched form (8 meshes x 32 tracers, FP32)")
        line["batched_form"] = batched_form(args.extra_steps)
        line["batched_form"]["per_cell_rate_over_single_mesh_fp32"] = \
            line["batched_form"]["updates_per_s"] / res["fp32"]["updates_per_s"]
        log(f"batched: {line['batched_form']['ms_per_step']:.1f} ms/step")
        log("learned limiter (1000x1000x64, 4 tracers)")
        line["learned_limiter"] = learned_form(api, args.extra_steps)
        log(f"learned: {line['learned_limiter']['ms_per_step']}")
    print(json.dumps(line))


if __name__ == "__main__":
    main()
