"""Batched scenes: every kernel of a substep covers all scenes of a batch
(scene table, concatenated slots and particles, per-scene CG).  A scene must
evolve in a batch as it does alone (the FP32 bar of one substep chain, 1e-5
relative L2: a batch reduces each scene's CG partials per block and then per
scene, a scene alone per work-group, so the FP64 sums may round apart)."""
import numpy as np
import pytest

import non_ficks_diffusion_amd as M

pytestmark = pytest.mark.gpu

SCENES = [("two_balls", 24, 0), ("blob", 32, 1), ("slab", 28, 2), ("avalanche", 32, 0), ("ref3d_test18", 20, 0),
          ("ref3d_test19", 32, 0)]  # the last: EOS fluid (every particle entry 0, an EOS material)


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


def load(scenes):
    out = []
    for name, counts, seed in scenes:
        d, x, v, m = M.scene_generate(name, counts, seed)
        out.append((d, x, v, m))
    return out


def run_alone(d, x, v, m, dts, material=None):
    s = M.Scene(d, len(m))
    s.upload(x, v, m, material=material)
    s.initialize()
    for dt in dts:
        s.substep(dt)
    r = s.download()
    st = s.stats()
    s.close()
    return r, st


def test_batch_matches_scenes_alone():
    data = load(SCENES)
    # per-scene dt sequences of different lengths: a scene with dt 0 idles
    dts = [[1e-3, 1e-3, 5e-4], [2e-3, 1e-3, 1e-3], [1e-3, 0.0, 1e-3], [5e-4, 5e-4, 5e-4], [2e-3, 2e-3, 0.0],
           [2e-3, 1e-3, 2e-3]]
    b = M.Batch([d for d, _, _, _ in data], [len(m) for _, _, _, m in data])
    for k, (d, x, v, m) in enumerate(data):
        b.upload(k, x, v, m)
    b.initialize()
    for step in range(3):
        b.substep([dts[k][step] for k in range(len(data))])
    for k, (d, x, v, m) in enumerate(data):
        alone, st_alone = run_alone(d, x, v, m, [t for t in dts[k] if t > 0])
        got = b.download(k)
        assert np.array_equal(got["ids"], alone["ids"]), k
        for f in ("x", "v", "F", "Jp", "volume"):
            assert rel(got[f], alone[f]) < 1e-5, (SCENES[k], f, rel(got[f], alone[f]))
        st = b.stats(k)
        assert st["live_particles"] == st_alone["live_particles"]
        assert st["active_blocks"] == st_alone["active_blocks"]


def test_idle_scene_is_unchanged():
    data = load(SCENES[:3])
    b = M.Batch([d for d, _, _, _ in data], [len(m) for _, _, _, m in data])
    for k, (d, x, v, m) in enumerate(data):
        b.upload(k, x, v, m)
    b.initialize()
    b.substep([1e-3, 1e-3, 1e-3])
    before = b.download(1)
    b.substep([1e-3, 0.0, 1e-3])
    after = b.download(1)
    for f in ("x", "v", "F", "Jp", "volume"):
        assert np.array_equal(before[f], after[f]), f


def test_batch_advance_to_time_matches_alone():
    data = load(SCENES[:3])
    target = 1.0 / 24.0
    b = M.Batch([d for d, _, _, _ in data], [len(m) for _, _, _, m in data])
    for k, (d, x, v, m) in enumerate(data):
        b.upload(k, x, v, m)
    b.initialize()
    targets = [target, 0.5 * target, target]
    t_out, steps = b.advance_to_time(0.0, targets)
    for k, (d, x, v, m) in enumerate(data):
        s = M.Scene(d, len(m))
        s.upload(x, v, m)
        s.initialize()
        t1, n1, _ = s.advance_to_time(0.0, targets[k])
        assert steps[k] == n1, (k, steps[k], n1)
        assert abs(t_out[k] - t1) < 1e-6
        got, alone = b.download(k), s.download()
        for f in ("x", "v", "F"):
            assert rel(got[f], alone[f]) < 1e-4, (k, f)
        s.close()


def test_batch_2d_matches_scenes_alone():
    """2D scenes (8^2 blocks) through the same batched kernels, with the
    two-material and EOS scenes among them (each scene its own table)."""
    scenes = [("disk2d", 64, 0, 0), ("disk2d", 48, 3, 800), ("ref2d_test18", 64, 0, 0), ("ref2d_test16", 64, 0, 0),
              ("ref2d_test19", 64, 0, 0)]
    data = [M.scene_generate(n, c, s, k) for n, c, s, k in scenes]
    mats = [M.scene_materials(n, c, s, k) for n, c, s, k in scenes]
    b = M.Batch([d for d, _, _, _ in data], [len(m) for _, _, _, m in data])
    for k, (d, x, v, m) in enumerate(data):
        b.upload(k, x, v, m, material=mats[k])
    b.initialize()
    dts = [1e-3, 1e-3, 2e-3, 2e-3, 1e-3]
    for _ in range(4):
        b.substep(dts)
    for k, (d, x, v, m) in enumerate(data):
        alone, _ = run_alone(d, x, v, m, [dts[k]] * 4, material=mats[k])
        got = b.download(k)
        assert np.array_equal(got["ids"], alone["ids"]), k
        assert np.array_equal(got["material"], mats[k][got["ids"]]), k
        for f in ("x", "v", "F", "Jp"):
            assert rel(got[f], alone[f]) < 1e-5, (scenes[k], f, rel(got[f], alone[f]))
    b.close()


@pytest.mark.parametrize("graph", [False, True])
def test_one_active_scene_takes_the_single_scene_solver(graph):
    # substeps where exactly one scene advances run the single-scene CG
    # kernels on it (launched, with the done poll, when it is large, even in
    # graph mode); every scene still matches itself alone
    specs = [("two_balls", 24, 0, 0), ("blob", 64, 1001, 70000), ("slab", 40, 1002, 8000)]
    data = [M.scene_generate(n, c, s, k) for n, c, s, k in specs]
    dts = [[1e-3, 0.0, 1e-3, 0.0], [0.0, 1e-3, 1e-3, 0.0], [0.0, 0.0, 1e-3, 1e-3]]
    b = M.Batch([d for d, _, _, _ in data], [len(m) for _, _, _, m in data])
    for k, (d, x, v, m) in enumerate(data):
        b.upload(k, x, v, m)
    b.initialize()
    if graph:
        b.use_graph(True)
    for step in range(4):
        b.substep([dts[k][step] for k in range(3)])
    for k, (d, x, v, m) in enumerate(data):
        alone, st_alone = run_alone(d, x, v, m, [t for t in dts[k] if t > 0])
        got = b.download(k)
        assert np.array_equal(got["ids"], alone["ids"]), k
        for f in ("x", "v", "F", "Jp", "volume"):
            assert rel(got[f], alone[f]) < 1e-5, (specs[k], f, rel(got[f], alone[f]))
        st = b.stats(k)
        assert st["live_particles"] == st_alone["live_particles"]
    b.close()
