"""Run-to-run reproducibility of the substep.

Binning ranks particles within a cell with atomics, whose arrival order
varies between runs; the reorder then sorts every cell by the particles'
previous positions (a stable sort by cell), so the particle layout, every
sum over a cell's particles and every CG reduction are the same in every
run.  Two handles fed the same input must therefore agree bit for bit after
each substep, CG iteration counts included, for one scene and for a batch.
"""
import numpy as np
import pytest

import non_ficks_diffusion_amd as M

pytestmark = pytest.mark.gpu

FIELDS = ("ids", "x", "v", "F", "Jp", "mass", "volume")


def assert_identical(a, b, what):
    for f in FIELDS:
        assert np.array_equal(a[f], b[f]), (what, f)


def test_repeated_single_scene_runs_are_bitwise_identical():
    d, x, v, m = M.scene_generate("blob", 32, 1001, 20000)
    runs = []
    for _ in range(2):
        s = M.Scene(d, len(m))
        s.upload(x, v, m)
        s.initialize()
        runs.append(s)
    for step in range(12):
        for s in runs:
            s.substep(1e-3)
        assert_identical(runs[1].download(), runs[0].download(), step)
        s0, s1 = runs[0].stats(), runs[1].stats()
        assert s0["cg_iterations"] == s1["cg_iterations"], step
        assert s0["cg_residual"] == s1["cg_residual"], step
    for s in runs:
        s.close()


def test_repeated_batch_runs_are_bitwise_identical():
    data = [M.scene_generate(n, c, seed, k) for n, c, seed, k in
            [("two_balls", 24, 0, 0), ("blob", 32, 1001, 20000), ("slab", 40, 1002, 8000)]]
    runs = []
    for _ in range(2):
        b = M.Batch([dd for dd, _, _, _ in data], [len(mm) for _, _, _, mm in data])
        for k, (_, x, v, m) in enumerate(data):
            b.upload(k, x, v, m)
        b.initialize()
        runs.append(b)
    for step in range(6):
        for b in runs:
            b.substep([1e-3, 5e-4, 1e-3])
        for k in range(len(data)):
            assert_identical(runs[1].download(k), runs[0].download(k), (step, k))
            assert runs[0].stats(k)["cg_iterations"] == runs[1].stats(k)["cg_iterations"], (step, k)
    for b in runs:
        b.close()
