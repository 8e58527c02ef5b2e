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
kernel (CG operator) measured with HIP events on the scene's stream, a
second window after the avalanche has developed, the CPU baseline (the
oracle, the C++/OpenMP restatement of the reference CPU path, timed on the
host cores on a bounded sample of the same workload) and the serving leg:
aggregate particle-substeps/s and p50/p99 frame latency of a bounded mixed
request stream, served with continuous batching and by the plain
one-scene-at-a-time loop.
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
# VALU wave-instructions per launch of the CG operator at this workload, from
# a rocprofv3 SQ_INSTS_VALU run, and MI355X's VALU issue peak without packed
# FP32 (256 CUs x 4 SIMDs x one wave64 instruction per 4 cycles x 2.4 GHz)
INSTRUCTIONS_FILE = os.path.join(ROOT, "profiles", "cg_force_avalanche_256_instructions.txt")
VALU_PEAK_G = 614.4
WORKLOAD = "avalanche"
COUNTS = 256
# substeps between the first timed window and the second (the developed avalanche)
DEVELOP = 1000
# the default run's bounded mixed request stream
SERVE_REQUESTS = 8
SERVE_FRAMES_MIN, SERVE_FRAMES_MAX = 1, 2


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


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(desc, x, v, m, dt, min_seconds=10.0, max_substeps=64):
    """Times the oracle (reference CPU path restated) on the host cores: one
    warm-up substep, then fixed-dt substeps until min_seconds have passed
    (a bounded sample of the same workload).  Each substep's CG iteration
    count is logged, the work-parity record beside the GPU's."""
    import oracle as O
    cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    o = O.Oracle(desc, 4, cores, O.SVD_FP64)  # float state, float CG dot products (as the reference)
    o.set_particles(x, v, m)
    o.initialize()
    o.advance_step(dt)  # warm-up substep
    t0 = time.perf_counter()
    substeps, el, its, timed = 0, 0.0, [], 0.0
    while substeps < max_substeps and (substeps == 0 or el < min_seconds):
        t1 = time.perf_counter()
        o.advance_step(dt)
        timed += time.perf_counter() - t1
        its.append(o.cg_stats()["cg_iterations"])
        substeps += 1
        el = time.perf_counter() - t0
    live = o.particles()["x"].shape[0]
    return {"value": live * substeps / timed, "unit": "particle-substeps/s", "cores": cores, "kind": "port",
            "cpu_model": cpu_model(), "cg_iterations": its,
            "sample": f"{WORKLOAD} {COUNTS}^3, {len(m)} particles, 1 warm-up + {substeps} fixed-dt substeps "
                      f"(dt {dt:.2e}, CG cap {desc.cg_max_iterations}), {timed:.1f} s"}


def measured_valu_insts(kernel="cg_force"):
    """<kernel>_valu_insts_per_launch from the committed counter summary, or None."""
    try:
        with open(INSTRUCTIONS_FILE) as f:
            for line in f:
                if line.startswith(kernel + "_valu_insts_per_launch:"):
                    return int(line.split(":", 1)[1].split()[0])
    except OSError:
        pass
    return None


def valu_roofline(kernel, ms):
    """VALU issue rate of one launch of `kernel` taking `ms`, against the issue peak."""
    insts = measured_valu_insts(kernel)
    if insts is None or not ms > 0:
        return None
    achieved = insts / (ms * 1e-3) / 1e9
    return {"kernel": kernel, "unit": "G wave-instructions/s", "insts_per_launch": insts,
            "achieved": round(achieved, 1), "peak": VALU_PEAK_G, "frac": round(achieved / VALU_PEAK_G, 4)}


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

    its = [b.stats(k)["cg_iterations"] for k in range(len(specs))]
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
    }
    if args.no_plain:
        b.close()
        print(json.dumps(out), flush=True)
        return

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
    for s in scenes:
        s.close()
    b.close()
    out.update({
        "plain_form": {"value": plain, "ms_per_step": el_p * 1e3 / args.steps,
                       "form": "one handle per scene, scenes advanced one after another"},
        "speedup_vs_plain": value / plain,
        "batch_vs_plain_rel_l2": {k: float(f"{v:.3e}") for k, v in err.items()},
        "checked_scenes": check,
    })
    print(json.dumps(out), flush=True)


def pct(a, q):
    return float(np.percentile(np.asarray(a, np.float64), q)) if len(a) else None


def latency_summary(first, inter):
    """Milliseconds: time to first frame (arrival to first frame, queueing
    included) and time between later frames of a request."""
    return {"first_frame_p50_ms": pct(first, 50), "first_frame_p99_ms": pct(first, 99),
            "frame_p50_ms": pct(inter, 50), "frame_p99_ms": pct(inter, 99),
            "frames": len(first) + len(inter)}


def fifo_replay(arrivals, service, frame_ends):
    """One server, first in first out: request k starts at max(its arrival,
    the end of request k-1) and its frames end frame_ends[k] seconds after
    it starts (its service does not depend on when it starts).  Returns
    (first-frame latencies ms, between-frame gaps ms, makespan s from the
    first arrival)."""
    first, inter = [], []
    end = -np.inf
    for a, s, fe in zip(arrivals, service, frame_ends):
        start = max(a, end)
        ft = [start + e for e in fe]
        first.append((ft[0] - a) * 1e3)
        inter.extend((np.diff(ft) * 1e3).tolist())
        end = start + s
    return first, inter, end - arrivals[0]


def bench_serve(args):
    """Serving form: a Poisson stream of snow-impact / avalanche requests,
    served by the plain loop (one scene at a time, in arrival order) and by
    the request server (continuous batching on one device arena), both
    against the same arrival clock.  The arrival rate defaults to 1.5 times
    the plain loop's capacity on the stream (requests / total service time)."""
    print(json.dumps(serve_stream(args)), flush=True)


def serve_stream(args):
    import non_ficks_diffusion_amd as M
    from non_ficks_diffusion_amd import workloads as W

    unit = W.serve_requests(args.requests, 1.0, frames_min=args.frames_min, frames_max=args.frames_max)
    data = [r.spec.generate() for r in unit]
    check = sorted({int(np.argmin([r.spec.n for r in unit])), int(np.argmax([r.spec.n for r in unit]))})
    if args.no_plain and args.rate is None:
        sys.exit("bench.py: --workload serve --no-plain needs --rate (the default rate is set from the plain loop)")

    plain = None
    if not args.no_plain:
        # plain form: one scene at a time, back to back.  A request's service
        # (handle, upload, frames) does not depend on when it starts, so its
        # waits against the arrival clock follow from first-in-first-out order.
        service, frame_ends, steps_p, kept = [], [], [], {}
        work_p = 0
        for k, r in enumerate(unit):
            t0 = time.perf_counter()
            d, x, v, m = data[k]
            s = M.Scene(d, len(m))
            s.upload(x, v, m)
            s.initialize()
            t, n_steps, ends = 0.0, 0, []
            for f in range(r.frames):
                target = float(np.float32(f + 1) / np.float32(d.frame_rate))
                t, n, _ = s.advance_to_time(t, target)
                n_steps += n
                s.synchronize()
                ends.append(time.perf_counter() - t0)
            work_p += r.spec.n * n_steps
            steps_p.append(n_steps)
            if k in check:
                kept[k] = s.download()
            s.close()
            service.append(time.perf_counter() - t0)
            frame_ends.append(ends)
        capacity = len(unit) / sum(service)
        plain = (service, frame_ends, steps_p, kept, work_p, capacity)
        print(f"plain loop: {len(unit)} requests served in {sum(service):.1f} s ({capacity:.3f} requests/s)",
              file=sys.stderr, flush=True)
    rate = args.rate if args.rate is not None else 1.5 * plain[5]
    reqs = W.serve_requests(args.requests, rate, frames_min=args.frames_min, frames_max=args.frames_max)
    assert [r.spec for r in reqs] == [r.spec for r in unit] and [r.frames for r in reqs] == [r.frames for r in unit]

    # batched form
    srv = M.Server(int(args.budget_gb * 1e9), max_scenes=args.max_scenes, max_counts=256, graph=args.graph,
                   keep_results=True, max_tick_substeps=args.tick_substeps, tune=args.autotune)
    arena = srv.arena()
    ids = [None] * len(reqs)
    offset = [0.0] * len(reqs)  # server clock minus this clock, at each submission
    nxt = 0
    work = 0
    ticks = 0
    bsteps = 0
    t0 = time.perf_counter()
    while True:
        now = time.perf_counter() - t0
        while nxt < len(reqs) and reqs[nxt].arrival <= now:
            d, x, v, m = data[nxt]
            ids[nxt] = srv.submit(d, x, v, m, reqs[nxt].frames)
            offset[nxt] = srv.poll(ids[nxt])["submitted"] - (time.perf_counter() - t0)
            nxt += 1
        if nxt == len(reqs) and all(srv.poll(i)["state"] == M.REQ_DONE for i in ids):
            break
        if all(srv.poll(ids[k])["state"] == M.REQ_DONE for k in range(nxt)):
            time.sleep(max(0.0, reqs[nxt].arrival - now))  # idle until the next arrival
            continue
        st = srv.tick()
        ticks += 1
        bsteps += st["substeps"]
        work += st["particle_substeps"]
    span_b = time.perf_counter() - t0 - reqs[0].arrival
    first, inter = [], []
    for k, i in enumerate(ids):
        ft = srv.frame_times(i)
        arrival_srv = reqs[k].arrival + offset[k]
        first.append((ft[0] - arrival_srv) * 1e3)
        inter.extend((np.diff(ft) * 1e3).tolist())
    served = {k: srv.download(ids[k]) for k in check}
    steps_b = [srv.poll(i)["substeps"] for i in ids]
    srv.close()
    lat_b = latency_summary(first, inter)
    ns = [r.spec.n for r in reqs]
    rate_note = (f"{rate:.3f} Hz (1.5x the plain loop's capacity)" if args.rate is None else f"{rate} Hz")
    out = {
        "metric": "particle-substeps/s and p50/p99 frame latency, mixed request stream",
        "value": work / span_b,
        "unit": "particle-substeps/s",
        "n_gpus": 1,
        "steps": bsteps,
        "warmup": 0,
        "ms_per_step": span_b * 1e3 / max(1, bsteps),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"{len(reqs)} Poisson requests at {rate_note}, snow-impact (two_balls) or "
                               f"avalanche, N log-uniform 10k-2M, grids 64^3/128^3/256^3, "
                               f"{args.frames_min}-{args.frames_max} frames at 24 fps",
                   "arrival_rate_hz": rate,
                   "particles_min": int(min(ns)), "particles_max": int(max(ns)),
                   "frames_total": int(sum(r.frames for r in reqs)), "hbm_budget_gb": args.budget_gb,
                   "arena": arena, "max_scenes": args.max_scenes, "graph": bool(args.graph),
                   "autotune": bool(args.autotune), "tick_substeps": args.tick_substeps, "ticks": ticks,
                   "request_substeps": int(sum(steps_b)),
                   "precision": "FP32 state, FP64 CG reductions and SVD", "parallelism": "single GPU"},
        "latency": lat_b,
        "makespan_s": span_b,
    }
    if plain is None:
        return out

    service, frame_ends, steps_p, kept, work_p, capacity = plain
    first, inter, span_p = fifo_replay([r.arrival for r in reqs], service, frame_ends)
    lat_p = latency_summary(first, inter)
    err = {"x": 0.0, "v": 0.0, "F": 0.0}
    for k in check:
        for f in err:
            err[f] = max(err[f], rel_l2(served[k][f], kept[k][f]))
    value, value_p = work / span_b, work_p / span_p
    out.update({
        "plain_form": {"value": value_p, "latency": lat_p, "makespan_s": span_p,
                       "capacity_hz": capacity, "service_s": sum(service),
                       "form": "one scene at a time in arrival order, a handle per request; served back to back, "
                               "waits against the arrival clock by first-in-first-out order"},
        "speedup_vs_plain": value / value_p,
        "substeps_equal": steps_b == steps_p,
        "served_vs_plain_rel_l2": {k: float(f"{v:.3e}") for k, v in err.items()},
        "checked_requests": check,
    })
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--counts", type=int, default=COUNTS)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-serve", action="store_true", help="skip the mixed-stream serving leg of the default run")
    ap.add_argument("--develop", type=int, default=DEVELOP,
                    help="substeps between the first timed window and the second (0: no second window)")
    ap.add_argument("--cpu-seconds", type=float, default=10.0,
                    help="CPU baseline: time fixed-dt oracle substeps for at least this long")
    ap.add_argument("--workload", choices=["avalanche", "batch", "serve"], default="avalanche",
                    help="avalanche: the 3D 1M-particle scene (default); batch: 256 scenes of uneven size, "
                         "batched against one scene per handle; serve: a Poisson request stream served with "
                         "continuous batching against the one-scene-at-a-time loop")
    ap.add_argument("--requests", type=int, default=24, help="serve: number of requests")
    ap.add_argument("--rate", type=float, default=None,
                    help="serve: Poisson arrival rate (requests/s); default 1.5x the plain loop's capacity")
    ap.add_argument("--frames-min", type=int, default=60, help="serve: frames per request, lower bound")
    ap.add_argument("--frames-max", type=int, default=240, help="serve: frames per request, upper bound")
    ap.add_argument("--budget-gb", type=float, default=200.0, help="serve: HBM budget of the arena")
    ap.add_argument("--max-scenes", type=int, default=32, help="serve: resident scenes at most")
    ap.add_argument("--tick-substeps", type=int, default=32,
                    help="serve: a tick also ends after this many substeps (new arrivals join at tick ends)")
    ap.add_argument("--scenes", type=int, default=256, help="batch: number of scenes")
    ap.add_argument("--graph", action="store_true", help="batch: replay a captured graph per substep")
    ap.add_argument("--autotune", action="store_true",
                    help="batch: pick launch shapes by the autotuner first; serve: re-tune as the resident set grows")
    ap.add_argument("--no-plain", action="store_true", help="batch, serve: skip the plain-form comparison")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run only a plain 1 GiB device-to-device copy (3 times), the reference kernel "
                         "for scaling the FETCH_SIZE / WRITE_SIZE counters")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: each scene runs on a single MI355X "
              f"(--gpus {args.gpus} is not supported)", file=sys.stderr)
        sys.exit(2)

    if args.copy_calibration:
        copy_calibration()
        return
    if args.workload == "batch":
        bench_batch(args)
        return
    if args.workload == "serve":
        bench_serve(args)
        return

    import non_ficks_diffusion_amd as M

    desc, x, v, m = M.scene_generate(WORKLOAD, args.counts, 0)
    n0 = len(m)
    s = M.Scene(desc, n0)
    s.upload(x, v, m)
    s.initialize()
    dt = cfl_dt(desc, s.max_particle_velocity())

    def timed_window(steps):
        """K substeps between two stream synchronisations: particle-substeps/s
        and ms per substep (live particles averaged over the window's ends)."""
        s.synchronize()
        live0 = s.stats()["live_particles"]
        t0 = time.perf_counter()
        for _ in range(steps):
            s.substep(dt)
        s.synchronize()
        el = time.perf_counter() - t0
        live1 = s.stats()["live_particles"]
        return steps * 0.5 * (live0 + live1) / el, el * 1e3 / steps

    for _ in range(args.warmup):
        s.substep(dt)
    value, ms_per_step = timed_window(args.steps)

    # per-kernel HIP-event timing pass (outside the timed region): the stats
    # of these substeps give the window's N, nodes and CG counts
    s.enable_timing(True)
    tsteps = 3
    stats = []
    for _ in range(tsteps):
        stats.append(s.substep(dt, stats=True))
    kt = s.kernel_times()
    s.enable_timing(False)
    N = stats[-1]["live_particles"]
    Nn = stats[-1]["active_blocks"] * 64
    I = float(np.mean([st["cg_iterations"] for st in stats]))
    Mul = I + 1.0
    cb = canonical_bytes(N, Nn, I, Mul)
    # dominant kernel: the CG operator, launched back to back on the state of
    # the last substep between two HIP events on the scene's stream
    avg_cg_ms = s.time_cg_operator(OPERATOR_REPS)
    achieved = cb["cg_force"] / (avg_cg_ms * 1e-3) / 1e9
    traffic = measured_traffic()
    breakdown = {k: round(msv / tsteps, 4) for k, (msv, n) in kt.items() if n}
    # P2G, G2P and the constitutive+force kernel timed the same way: launched
    # back to back on the last substep's state (restored afterwards)
    live_ms = {"p2g": s.time_stage(M.STAGE_P2G, OPERATOR_REPS),
               "g2p": s.time_stage(M.STAGE_G2P, OPERATOR_REPS),
               "constitutive_force": s.time_stage(M.STAGE_CONSTITUTIVE, OPERATOR_REPS)}
    per_kernel_gbs = {k: round(cb[k] / (live_ms[k] * 1e-3) / 1e9, 1) for k in live_ms}
    per_kernel_gbs["cg_force"] = round(achieved, 1)
    per_kernel_frac = {k: round(v / HBM_PEAK_GBS, 4) for k, v in per_kernel_gbs.items()}
    # measured roof beside the 8 TB/s peak: a streaming 1 GiB copy
    copy_gbs = M.measure_copy_bandwidth(1 << 30, 10)
    per_kernel_copy_frac = {k: round(v / copy_gbs, 4) for k, v in per_kernel_gbs.items()}

    # the developed avalanche: --develop more substeps at the same dt
    # (untimed), then the same K substeps timed again; the CG counts are those
    # of three substeps after the window
    developed = None
    if args.develop > 0:
        for _ in range(args.develop):
            s.substep(dt)
        value_dev, ms_dev = timed_window(args.steps)
        st_dev = [s.substep(dt, stats=True) for _ in range(tsteps)]
        I_dev = float(np.mean([st["cg_iterations"] for st in st_dev]))
        cb_dev = canonical_bytes(st_dev[-1]["live_particles"], st_dev[-1]["active_blocks"] * 64, I_dev, I_dev + 1.0)
        done = args.warmup + args.steps + tsteps + args.develop
        developed = {"after_substeps": done, "sim_time_s": round(done * dt, 4),
                     "value": value_dev, "ms_per_step": ms_dev, "steps": args.steps,
                     "particles": int(st_dev[-1]["live_particles"]), "active_blocks": int(st_dev[-1]["active_blocks"]),
                     "cg_iterations": [int(st["cg_iterations"]) for st in st_dev],
                     "substep_achieved_GBs": round(cb_dev["substep"] / (ms_dev * 1e-3) / 1e9, 1)}

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
                   "dt": dt,
                   "dt_rule": "fixed-dt mode: the first substep's dt, clamp(cfl dx / max(vmax, 1e-2), min_dt, max_dt) "
                              "with max_dt the elastic bound dx / c_p (c_p = sqrt((lambda0 + 2 mu0) / rho)), not "
                              "the 5e-3 cap: from rest it is the elastic bound",
                   "cg_iterations": round(I, 2), "cg_multiplies": round(Mul, 2),
                   "cg_tolerance": float(desc.cg_tolerance), "cg_max_iterations": int(desc.cg_max_iterations),
                   "precision": "FP32 state, FP64 CG reductions and SVD", "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "kernel": "cg_force", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "avg_launch_ms": round(avg_cg_ms, 5), "launches_timed": OPERATOR_REPS,
                     "algorithmic_bytes_per_launch": int(cb["cg_force"])},
        # the operator's own bound: VALU issue (counter-measured instructions per launch
        # over the live launch time, against the issue peak)
        "valu_roofline": valu_roofline("cg_force", avg_cg_ms),
        # the FP64-SVD constitutive update and explicit force: VALU too
        "valu_roofline_constitutive": valu_roofline("constitutive_force", live_ms["constitutive_force"]),
        # the whole substep: canonical bytes over the timed window's ms per substep
        "substep_roofline": {"algorithmic_bytes": int(cb["substep"]), "ms": round(ms_per_step, 4),
                             "achieved_GBs": round(cb["substep"] / (ms_per_step * 1e-3) / 1e9, 1),
                             "frac": round(cb["substep"] / (ms_per_step * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)},
        "kernel_GBs": per_kernel_gbs,
        "kernel_frac_of_hbm_peak": per_kernel_frac,
        "copy_roof": {"GBs": round(copy_gbs, 1), "bytes": 1 << 30, "launches": 10,
                      "form": "streaming device-to-device copy, read plus written bytes"},
        "kernel_frac_of_copy_roof": per_kernel_copy_frac,
        # each kernel class between its own HIP events over three substeps: the
        # events serialise the launches, so the sum exceeds ms_per_step
        "event_timed_ms_per_substep": breakdown,
        "kernel_live_ms": {k: round(v, 5) for k, v in {**live_ms, "cg_force": avg_cg_ms}.items()},
        "developed": developed,
    }
    s.close()
    if not args.no_cpu:
        cpu = cpu_baseline(desc, x, v, m, dt, args.cpu_seconds)
        out["cpu_baseline"] = cpu
        out["gpu_over_cpu"] = round(value / cpu["value"], 1)
        gpu_its = [int(st["cg_iterations"]) for st in stats]
        # per CG operator application (iterations + 1 per substep): the same work unit on both sides
        per_app = (value * (np.mean(gpu_its) + 1.0)) / (cpu["value"] * (np.mean(cpu["cg_iterations"]) + 1.0))
        out["work_parity"] = {"gpu_cg_iterations": gpu_its, "cpu_cg_iterations": cpu["cg_iterations"],
                              "gpu_over_cpu_per_operator_application": round(float(per_app), 1),
                              "note": "same scene, dt and CG cap; the CPU sample runs the first substeps from "
                                      "the start, the GPU figures are from its timed window"}
    if not args.no_serve:
        out["serving"] = serving_leg(args)
    print(json.dumps(out), flush=True)


def serving_leg(args):
    """The mixed request stream on a bounded sample: SERVE_REQUESTS Poisson
    requests of SERVE_FRAMES_MIN-SERVE_FRAMES_MAX frames, served by the request
    server (graph replays, autotuned launch shapes) and by the plain
    one-scene-at-a-time loop against the same arrival clock."""
    sa = argparse.Namespace(requests=SERVE_REQUESTS, rate=None, frames_min=SERVE_FRAMES_MIN,
                            frames_max=SERVE_FRAMES_MAX, budget_gb=200.0, max_scenes=32, graph=True,
                            autotune=True, tick_substeps=32, no_plain=False)
    r = serve_stream(sa)
    lat, lat_p = r["latency"], r["plain_form"]["latency"]
    return {"workload": r["config"]["workload"], "requests": SERVE_REQUESTS,
            "frames_total": r["config"]["frames_total"], "value": r["value"], "unit": "particle-substeps/s",
            "frame_latency_p50_ms": lat["frame_p50_ms"], "frame_latency_p99_ms": lat["frame_p99_ms"],
            "first_frame_p50_ms": lat["first_frame_p50_ms"], "first_frame_p99_ms": lat["first_frame_p99_ms"],
            "plain_value": r["plain_form"]["value"],
            "plain_frame_latency_p50_ms": lat_p["frame_p50_ms"], "plain_frame_latency_p99_ms": lat_p["frame_p99_ms"],
            "plain_first_frame_p50_ms": lat_p["first_frame_p50_ms"],
            "plain_first_frame_p99_ms": lat_p["first_frame_p99_ms"],
            "speedup_vs_plain": r["speedup_vs_plain"], "substeps_equal": r["substeps_equal"],
            "served_vs_plain_rel_l2": r["served_vs_plain_rel_l2"], "makespan_s": r["makespan_s"],
            "plain_makespan_s": r["plain_form"]["makespan_s"]}


if __name__ == "__main__":
    main()
