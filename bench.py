#!/usr/bin/env python3
"""Throughput benchmark of the MI355X MPM snow substep.

Workload: the 3D sparse avalanche -- a 0.6 x 0.1 x 0.125 snow slab (8
particles per cell, about 1.0M particles) on a 30-degree incline in a 256^3
grid, snow material (rho 400, E 1.4e5, nu 0.2, theta_c 2.5e-2,
theta_s 7.5e-3, xi 10), CG tolerance 1e-6 / at most 50 iterations.  A step
is one substep (MPM_Driver::Advance_Step) in fixed-dt mode: the dt of the
first CFL substep, kept for every timed substep.  Particle state, grid and
CG vectors stay resident in HBM; the timed region contains only kernel work.

Prints one JSON line: particle-substeps/s, the roofline of the dominant
kernel (CG operator) measured with HIP events on the scene's stream, and the
CPU baseline (the oracle, the C++/OpenMP restatement of the reference CPU
path, timed on the host cores on a bounded sample of the same workload).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
OPERATOR_REPS = 20
# HBM bytes per launch of the CG operator at this workload, from rocprofv3
# FETCH_SIZE and WRITE_SIZE runs scaled by a copy kernel of known size
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "cg_force_avalanche_256_traffic.txt")
WORKLOAD = "avalanche"
COUNTS = 256


def canonical_bytes(N, Nn, I, M):
    """Algorithmic HBM bytes per substep: every particle and grid field read or written once (FP32 = 4 B)."""
    return {
        "substep": N * (320 + 60 * M) + Nn * (200 + 64 * M + 164 * I),
        "cg_force": N * 60 + Nn * 24,      # per operator application: particles + gather s + scatter q
        "p2g": N * 32 + Nn * 16,
        "g2p": N * 124 + Nn * 24,
        "constitutive_force": N * (80 + 64) + Nn * 12,
    }


def cfl_dt(desc, vmax):
    dx = 1.0 / desc.counts
    return float(max(desc.min_dt, min(desc.max_dt, desc.cfl * dx / max(vmax, 1e-2))))


def cpu_baseline(desc, x, v, m, dt, substeps=2):
    """Times the oracle (reference CPU path restated) on the host cores."""
    import oracle as O
    cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    o = O.Oracle(desc, 4, cores, O.SVD_FP64)  # float state, float CG dot products (as the reference)
    o.set_particles(x, v, m)
    o.initialize()
    o.advance_step(dt)  # warm-up substep
    t0 = time.perf_counter()
    for _ in range(substeps):
        o.advance_step(dt)
    el = time.perf_counter() - t0
    live = o.particles()["x"].shape[0]
    return {"value": live * substeps / el, "unit": "particle-substeps/s", "cores": cores, "kind": "port",
            "sample": f"{WORKLOAD} {COUNTS}^3, {len(m)} particles, 1 warm-up + {substeps} fixed-dt substeps "
                      f"(dt {dt:.2e}, CG cap {desc.cg_max_iterations}), {el:.1f} s"}


def measured_traffic():
    """traffic_bytes_per_launch from the committed counter summary, or None."""
    try:
        with open(TRAFFIC_FILE) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:"):
                    return int(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def copy_calibration():
    """A copy kernel of known size: 1 GiB read + 1 GiB written per launch."""
    import torch
    n = (1 << 30) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read_per_launch": n * 4, "copy_bytes_written_per_launch": n * 4, "launches": 3}),
          flush=True)


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


def bench_batch(args):
    """Batched form: 256 scenes of uneven size in one batch, against the
    plain form (one scene per handle, advanced one after another) on the same
    GPU with the same per-scene dt and CG settings."""
    import non_ficks_diffusion_amd as M
    from non_ficks_diffusion_amd import workloads as W

    specs = W.batch_scenes(args.scenes)
    data = [sp.generate() for sp in specs]
    descs = [d for d, _, _, _ in data]
    caps = [len(m) for _, _, _, m in data]
    b = M.Batch(descs, caps)
    for k, (_, x, v, m) in enumerate(data):
        b.upload(k, x, v, m)
    b.initialize()
    vmax = b.max_velocity()
    dts = np.array([cfl_dt(d, vmax[k]) for k, d in enumerate(descs)], np.float32)
    tune = None
    if args.autotune:
        tune = b.autotune(dts)
    if args.graph:
        b.use_graph(True)

    def run_batch(steps):
        for _ in range(steps):
            b.substep(dts)
        b.synchronize()

    run_batch(args.warmup)
    t0 = time.perf_counter()
    run_batch(args.steps)
    el_b = time.perf_counter() - t0
    live = sum(b.stats(k)["live_particles"] for k in range(len(specs)))
    value = live * args.steps / el_b

    # plain form: one handle per scene, each substep its own launches
    scenes = []
    for (d, x, v, m) in data:
        s = M.Scene(d, len(m))
        s.upload(x, v, m)
        s.initialize()
        scenes.append(s)

    def run_plain(steps):
        for _ in range(steps):
            for k, s in enumerate(scenes):
                s.substep(float(dts[k]))
        for s in scenes:
            s.synchronize()

    run_plain(args.warmup)
    t0 = time.perf_counter()
    run_plain(args.steps)
    el_p = time.perf_counter() - t0
    live_p = sum(s.stats()["live_particles"] for s in scenes)
    plain = live_p * args.steps / el_p

    # equal accuracy: both forms ran the same substeps with the same dt
    check = sorted({0, int(np.argmin(caps)), int(np.argmax(caps)), len(specs) // 2})
    err = {"x": 0.0, "v": 0.0, "F": 0.0}
    for k in check:
        got, ref = b.download(k), scenes[k].download()
        for f in err:
            err[f] = max(err[f], rel_l2(got[f], ref[f]))
    its = [b.stats(k)["cg_iterations"] for k in range(len(specs))]
    for s in scenes:
        s.close()
    b.close()
    n_arr = np.array(caps)
    out = {
        "metric": "particle-substeps/s, 256-scene batch",
        "value": value,
        "unit": "particle-substeps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": el_b * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"batch of {len(specs)} 3D snow blobs/slabs, seeds 1000+s, grids 32^3-128^3, "
                               f"N log-uniform 2k-200k",
                   "particles": int(n_arr.sum()), "n_min": int(n_arr.min()), "n_max": int(n_arr.max()),
                   "cg_iterations_max": int(max(its)), "cg_iterations_median": float(np.median(its)),
                   "graph": bool(args.graph), "autotune": tune,
                   "precision": "FP32 state, FP64 CG reductions and SVD", "parallelism": "single GPU"},
        "plain_form": {"value": plain, "ms_per_step": el_p * 1e3 / args.steps,
                       "form": "one handle per scene, scenes advanced one after another"},
        "speedup_vs_plain": value / plain,
        "batch_vs_plain_rel_l2": {k: float(f"{v:.3e}") for k, v in err.items()},
        "checked_scenes": check,
    }
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--counts", type=int, default=COUNTS)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-substeps", type=int, default=2)
    ap.add_argument("--workload", choices=["avalanche", "batch"], default="avalanche",
                    help="avalanche: the 3D 1M-particle scene (default); batch: 256 scenes of uneven size, "
                         "batched against one scene per handle")
    ap.add_argument("--scenes", type=int, default=256, help="batch: number of scenes")
    ap.add_argument("--graph", action="store_true", help="batch: replay a captured graph per substep")
    ap.add_argument("--autotune", action="store_true", help="batch: pick launch shapes by the autotuner first")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run only a plain 1 GiB device-to-device copy (3 times), the reference kernel "
                         "for scaling the FETCH_SIZE / WRITE_SIZE counters")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build runs each scene on a single MI355X "
              f"(--gpus {args.gpus} is not supported)", file=sys.stderr)
        sys.exit(2)

    if args.copy_calibration:
        copy_calibration()
        return
    if args.workload == "batch":
        bench_batch(args)
        return

    import non_ficks_diffusion_amd as M

    desc, x, v, m = M.scene_generate(WORKLOAD, args.counts, 0)
    n0 = len(m)
    s = M.Scene(desc, n0)
    s.upload(x, v, m)
    s.initialize()
    dt = cfl_dt(desc, s.max_particle_velocity())

    for _ in range(args.warmup):
        s.substep(dt)
    s.synchronize()
    live0 = s.stats()["live_particles"]
    # timed region: K substeps, synchronised on both sides
    t0 = time.perf_counter()
    live_sum = 0
    for _ in range(args.steps):
        s.substep(dt)
    s.synchronize()
    el = time.perf_counter() - t0
    st = s.stats()
    # live particles before the timed region plus after; removal is rare and
    # monotone, so the average bounds the per-substep count
    live_sum = args.steps * 0.5 * (live0 + st["live_particles"])
    value = live_sum / el
    ms_per_step = el * 1e3 / args.steps

    # per-kernel HIP-event timing pass (outside the timed region)
    s.enable_timing(True)
    tsteps = 3
    stats = []
    for _ in range(tsteps):
        stats.append(s.substep(dt, stats=True))
    kt = s.kernel_times()
    s.enable_timing(False)
    N = stats[-1]["live_particles"]
    Nn = stats[-1]["active_blocks"] * 64
    I = stats[-1]["cg_iterations"]
    Mul = stats[-1]["cg_multiplies"]
    cb = canonical_bytes(N, Nn, I, Mul)
    # dominant kernel: the CG operator, launched back to back on the state of
    # the last substep between two HIP events on the scene's stream
    avg_cg_ms = s.time_cg_operator(OPERATOR_REPS)
    achieved = cb["cg_force"] / (avg_cg_ms * 1e-3) / 1e9
    traffic = measured_traffic()
    breakdown = {k: round(msv / tsteps, 4) for k, (msv, n) in kt.items() if n}
    sub_ms = sum(msv for msv, _ in kt.values()) / tsteps
    per_kernel_gbs = {}
    for k in ("p2g", "g2p", "constitutive_force"):
        msv, n = kt[k]
        if n:
            per_kernel_gbs[k] = round(cb[k] / (msv / n * 1e-3) / 1e9, 1)
    per_kernel_gbs["cg_force"] = round(achieved, 1)
    per_kernel_frac = {k: round(v / HBM_PEAK_GBS, 4) for k, v in per_kernel_gbs.items()}

    out = {
        "metric": "particle-substeps/s, 3D 1M snow",
        "value": value,
        "unit": "particle-substeps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"{WORKLOAD} (3D sparse avalanche, 30-degree incline)", "grid": f"{args.counts}^3",
                   "particles": int(N), "active_blocks": int(stats[-1]["active_blocks"]),
                   "active_block_fraction": round(stats[-1]["active_blocks"] / ((args.counts // 4) ** 3), 4),
                   "dt": dt, "cg_iterations": int(I), "cg_multiplies": int(Mul),
                   "cg_tolerance": float(desc.cg_tolerance), "cg_max_iterations": int(desc.cg_max_iterations),
                   "precision": "FP32 state, FP64 CG reductions and SVD", "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "kernel": "cg_force", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "avg_launch_ms": round(avg_cg_ms, 5), "launches_timed": OPERATOR_REPS,
                     "algorithmic_bytes_per_launch": int(cb["cg_force"])},
        "substep_roofline": {"algorithmic_bytes": int(cb["substep"]), "kernel_ms": round(sub_ms, 4),
                             "achieved_GBs": round(cb["substep"] / (sub_ms * 1e-3) / 1e9, 1)},
        "kernel_GBs": per_kernel_gbs,
        "kernel_frac_of_hbm_peak": per_kernel_frac,
        "kernel_ms_per_substep": breakdown,
    }
    s.close()
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(desc, x, v, m, dt, args.cpu_substeps)
        out["gpu_over_cpu"] = round(value / out["cpu_baseline"]["value"], 1)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
