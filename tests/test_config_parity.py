"""GPU parity of MPM_Driver's dt rule, the Project reading and the workload sizes
against the CPU oracle, through the C ABI.

* The dt rule and max speed (MPM_Driver.cpp:55-80, MPM_Example.cpp:375-387)
  on a moving scene, with the host loop and the device frame loop.
* Project_Helper's "zero inside planes" reading (project_mode 0,
  Implicit_Force_Helper/Project_Helper.h:30-45), stage by stage.
* The two-ball impact at its stated size (128^3, 242,863 particles): one
  substep against the FP64 oracle, 100 substeps against the FP32 oracle.
* The full 256-scene batch (graph replays, autotuned launch shapes): checked
  scenes against the oracle, each run alone on the CPU.
* A served request mix holding a 256^3 avalanche of 1M particles: every
  request against the oracle advanced frame by frame to the same frame ends.
* The viewer's block flags (MPM_Example.cpp:1061-1062) against the oracle's
  Interior flags.

Bars: 1e-5 relative L2 for one or two substeps; the substep-chain bar of the
target-time tests (1e-4) when dts follow the state; 1e-3 after 100 substeps.
"""
import os

import numpy as np
import pytest

import non_ficks_diffusion_amd as M
import oracle as O
from non_ficks_diffusion_amd import workloads as W

pytestmark = pytest.mark.gpu

MATCHED = O.DOT_FP64 | O.SVD_FP64
THREADS = max(1, min(16, len(os.sched_getaffinity(0))))


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


def rel_m(a, b, mass):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    w = np.asarray(mass, np.float64)[..., None]
    return float(np.sqrt((w * (a - b) ** 2).sum()) / max(np.sqrt((w * b ** 2).sum()), 1e-30))


def compare(g, r, bar, what, fields=("x", "v", "F", "Jp")):
    assert np.array_equal(g["ids"], r["ids"]), what
    errs = {k: rel(g[k], r[k]) for k in fields}
    print(what, {k: f"{e:.2e}" for k, e in errs.items()}, flush=True)
    for k, e in errs.items():
        assert e < bar, (what, k, e)


def first_dt(d, v):
    """The reference's first dt (MPM_Driver.cpp:67) in float32 from the
    uploaded velocities."""
    vmax = np.float32(np.sqrt(np.max((v.astype(np.float32) ** 2).sum(1)))) if len(v) else np.float32(0)
    dx = np.float32(1.0) / np.float32(d.counts)
    dt = np.float32(d.cfl) * dx / max(vmax, np.float32(1e-2))
    return float(max(np.float32(d.min_dt), min(np.float32(d.max_dt), dt)))


# --------------------------------------------------------------------------- A2
@pytest.mark.parametrize("graph", [False, True])
def test_dt_rule_and_max_speed_on_a_moving_scene(graph):
    """Two balls at 3 m/s each, advanced over two frames that contain the
    impact (so the max speed, and with it the CFL dt, changes from substep
    to substep): the same substep count and end time as the oracle's
    Advance_To_Target_Time, the max speed of the oracle before and after,
    the state to the substep-chain bar."""
    d, x, v, m = M.scene_generate("two_balls", 32, 0)
    s = M.Scene(d, len(m))
    s.upload(x, v, m)
    s.initialize()
    if graph:
        s.use_graph(True)
    o = O.Oracle(d, 4, THREADS, MATCHED)
    o.set_particles(x, v, m)
    o.initialize()
    assert abs(s.max_particle_velocity() - o.max_particle_velocity()) <= 1e-6 * o.max_particle_velocity()
    t_g = t_o = 0.0
    for frame in (1, 2):
        target = float(np.float32(frame) / np.float32(24.0))
        t_g, n_g, _ = s.advance_to_time(t_g, target)
        t_o, n_o = o.advance_to_target_time(t_o, target)
        print(f"frame {frame}: gpu {n_g} substeps to {t_g!r}, oracle {n_o} to {t_o!r}", flush=True)
        assert n_g == n_o, (frame, n_g, n_o)
        assert abs(t_g - t_o) <= 1e-6 * t_o
        vg, vo = s.max_particle_velocity(), o.max_particle_velocity()
        assert abs(vg - vo) <= 1e-4 * vo, (vg, vo)
    compare(s.download(), o.particles(), 1e-4, f"two balls 32^3 two frames (graph={graph})")
    s.close()


# --------------------------------------------------------------------------- A21
def test_project_zero_inside_planes_stage_by_stage():
    """project_mode 0: the CG vectors vanish at nodes inside a plane, so those
    nodes keep their post-collision v* through the solve.  Every stage of
    Advance_Step against the FP64 oracle at 1e-5; the reading changes the
    solve (the same scene with the no-op reading ends elsewhere)."""
    from test_gpu_parity import run_both_stages

    d, x, v, m = M.scene_generate("avalanche", 64, 0)
    d.cg_max_iterations = 20
    d.project_mode = 0
    s = M.Scene(d, len(m))
    s.upload(x, v, m)
    o = O.Oracle(d, 8, THREADS, MATCHED)
    o.set_particles(x, v, m)
    errs = run_both_stages(s, o, 1e-3)
    print("project_mode 0", {k: f"{e:.1e}" for k, e in errs.items()}, flush=True)
    for k, e in errs.items():
        assert e < 1e-5, (k, e)
    zeroed = o.particles()
    d1 = d.copy()
    d1.project_mode = 1
    o1 = O.Oracle(d1, 8, THREADS, MATCHED)
    o1.set_particles(x, v, m)
    for st in range(M.STAGE_BIN, M.STAGE_G2P + 1):
        o1.run_stage(st, 1e-3)
    assert rel(zeroed["v"], o1.particles()["v"]) > 1e-6
    s.close()


def test_project_zero_inside_planes_one_substep():
    """The fused substep with project_mode 0 against the FP64 oracle."""
    d, x, v, m = M.scene_generate("two_balls", 32, 0)
    d.project_mode = 0
    s = M.Scene(d, len(m))
    s.upload(x, v, m)
    s.initialize()
    o = O.Oracle(d, 8, THREADS, MATCHED)
    o.set_particles(x, v, m)
    o.initialize()
    for _ in range(3):
        s.substep(1e-3)
        o.advance_step(1e-3)
    compare(s.download(), o.particles(), 1e-5, "two balls 32^3 project_mode 0, 3 substeps")
    s.close()


# --------------------------------------------------------------------------- config 1
def test_two_balls_128_one_substep_against_fp64_oracle():
    d, x, v, m = M.scene_generate("two_balls", 128, 0)
    assert len(m) > 240_000
    dt = first_dt(d, v)
    s = M.Scene(d, len(m))
    s.upload(x, v, m)
    s.initialize()
    st = s.substep(dt, stats=True)
    g = s.download()
    s.close()
    o = O.Oracle(d, 8, THREADS, MATCHED)
    o.set_particles(x, v, m)
    o.initialize()
    o.advance_step(dt)
    print("two balls 128^3: gpu cg", st["cg_iterations"], "fp64 oracle cg", o.cg_stats()["cg_iterations"])
    compare(g, o.particles(), 1e-5, f"two balls 128^3 one substep (dt {dt:.3e}) against FP64")


def test_two_balls_128_hundred_substeps():
    """100 substeps at the elastic dt cap (3.96e-4): the balls meet near
    substep 68, so the impact is inside the chain."""
    d, x, v, m = M.scene_generate("two_balls", 128, 0)
    dt = float(np.float32(d.max_dt))
    s = M.Scene(d, len(m))
    s.upload(x, v, m)
    s.initialize()
    o = O.Oracle(d, 4, THREADS, MATCHED)
    o.set_particles(x, v, m)
    o.initialize()
    for k in range(100):
        s.substep(dt)
        o.advance_step(dt)
    g, r = s.download(), o.particles()
    s.close()
    compare(g, r, 1e-3, "two balls 128^3 after 100 substeps against FP32")


# --------------------------------------------------------------------------- config 3
def test_full_batch_against_the_oracle():
    """The 256-scene batch (2k-200k particles each, 32^3-128^3 grids), launch
    shapes from the autotuner and every substep a graph replay, two substeps
    with each scene's own first dt.  Six scenes (the smallest, the largest
    and four drawn by a seed) against the oracle run alone."""
    specs = W.batch_scenes()
    data = [sp.generate() for sp in specs]
    b = M.Batch([d for d, _, _, _ in data], [len(m) for _, _, _, m in data])
    for k, (d, x, v, m) in enumerate(data):
        b.upload(k, x, v, m)
    b.initialize()
    dts = np.array([first_dt(d, v) for d, _, v, _ in data], np.float32)
    b.autotune(dts)
    b.use_graph(True)
    for _ in range(2):
        b.substep(dts)
    ns = np.array([len(m) for _, _, _, m in data])
    pick = [int(np.argmin(ns)), int(np.argmax(ns))]
    pick += [int(k) for k in np.random.default_rng(99).choice(len(data), 6, replace=False) if k not in pick][:4]
    print("total particles", int(ns.sum()), "checked scenes", pick, flush=True)
    for k in pick:
        d, x, v, m = data[k]
        o = O.Oracle(d, 4, THREADS, MATCHED)
        o.set_particles(x, v, m)
        o.initialize()
        for _ in range(2):
            o.advance_step(float(dts[k]))
        compare(b.download(k), o.particles(), 1e-5, f"batch scene {k} ({specs[k].name} {specs[k].counts}^3, "
                                                     f"{specs[k].n} particles)")
    b.close()


# --------------------------------------------------------------------------- config 4
SERVED = [("two_balls", 64, 5001, 20_000), ("avalanche", 128, 5002, 90_000), ("avalanche", 256, 5003, 1_000_000)]
SERVE_FPS = 2400.0   # frames of a few substeps, so the oracle follows every one
SERVE_FRAMES = 2


@pytest.fixture(scope="module")
def served_oracle():
    out = []
    for name, c, seed, n in SERVED:
        d, x, v, m = M.scene_generate(name, c, seed, n)
        d.frame_rate = SERVE_FPS
        o = O.Oracle(d, 4, THREADS, MATCHED)
        o.set_particles(x, v, m)
        o.initialize()
        t, steps = 0.0, 0
        for f in range(SERVE_FRAMES):
            t, k = o.advance_to_target_time(t, float(np.float32(f + 1) / np.float32(SERVE_FPS)))
            steps += k
        out.append(((d, x, v, m), o.particles(), steps))
    return out


@pytest.mark.parametrize("graph", [False, True])
def test_served_mix_with_a_256_request_against_the_oracle(served_oracle, graph):
    """Requests of 20k (64^3), 90k (128^3) and 1M (256^3) particles; the first
    two are admitted at once, the 256^3 one arrives while they run.  Each
    finished request equals the oracle advanced to the same frame ends, with
    the same substep count."""
    srv = M.Server(8 << 30, max_scenes=4, max_counts=256, graph=graph, keep_results=True)
    ids = []
    for (d, x, v, m), _, _ in served_oracle[:2]:
        ids.append(srv.submit(d, x, v, m, SERVE_FRAMES))
    srv.tick()
    (d, x, v, m), _, _ = served_oracle[2]
    ids.append(srv.submit(d, x, v, m, SERVE_FRAMES))
    ticks = 1
    while any(srv.poll(i)["state"] != M.REQ_DONE for i in ids):
        srv.tick()
        ticks += 1
        assert ticks < 20
    for rid, ((d, x, v, m), r, steps) in zip(ids, served_oracle):
        st = srv.poll(rid)
        assert st["substeps"] == steps, (d.counts, st["substeps"], steps)
        compare(srv.download(rid), r, 1e-4, f"served {d.counts}^3 {len(m)} particles (graph={graph})")
    srv.close()


# --------------------------------------------------------------------------- F3
@pytest.mark.parametrize("name,counts", [("two_balls", 32), ("avalanche", 64), ("disk2d", 64)])
def test_block_flags_against_the_oracle(name, counts):
    """The viewer's per-block flags (mpm_read_block_flags) after P2G against
    the FP32 oracle's Interior flags (Flag_Setup_Helper.h:25-35): every flag
    inside the domain equal, nodes outside carry none, and every Interior
    node of the oracle lies in an active block."""
    d, x, v, m = M.scene_generate(name, counts, 0)
    s = M.Scene(d, len(m))
    s.upload(x, v, m)
    s.run_stage(M.STAGE_BIN)
    s.run_stage(M.STAGE_P2G)
    o = O.Oracle(d, 4, THREADS, MATCHED)
    o.set_particles(x, v, m)
    o.run_stage(M.STAGE_BIN)
    o.run_stage(M.STAGE_P2G)
    ref = o.flags() & M.MPM_FLAG_INTERIOR
    origin, flags = s.read_block_flags()
    covered = np.zeros(ref.shape, bool)
    for b in range(len(origin)):
        for l in range(64):
            lc = [l >> 4, (l >> 2) & 3, l & 3] if d.dim == 3 else [l >> 3, l & 7]
            node = [int(origin[b][a]) + lc[a] for a in range(d.dim)]
            f = int(flags[b, l]) & M.MPM_FLAG_INTERIOR
            if all(-1 <= i <= counts for i in node):
                idx = tuple(i + 1 for i in node)
                assert f == ref[idx], (node, f, ref[idx])
                covered[idx] = True
            else:
                assert f == 0, node
    assert np.all(covered[ref != 0])
    s.close()
