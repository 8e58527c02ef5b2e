#!/usr/bin/env python3
"""Writes the small regression fixtures of tests/golden/ from the FP64 oracle.

Each fixture holds a seeded scene's particle inputs and the state after a
fixed number of fixed-dt substeps (FP64 oracle, one thread, deterministic).
They pin the oracle against regressions and give the GPU tests a stored
expected output.  Run from the repository root:  python tests/golden/make_fixtures.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import non_ficks_diffusion_amd as M  # noqa: E402
import oracle as O  # noqa: E402

CASES = [
    # name, scene, counts, dt, substeps
    ("disk2d_64_5steps", "disk2d", 64, 5e-4, 5),
    ("two_balls_16_3steps", "two_balls", 16, 1e-3, 3),
    ("avalanche_32_3steps", "avalanche", 32, 1e-3, 3),
    # the reference's 2D test 16 (two materials) and test 19 (EOS fluid)
    ("ref2d_test16_64_5steps", "ref2d_test16", 64, 2e-3, 5),
    ("ref2d_test19_64_5steps", "ref2d_test19", 64, 2e-3, 5),
]


def make(name, scene, counts, dt, steps):
    d, x, v, m = M.scene_generate(scene, counts, 0)
    mat = M.scene_materials(scene, counts, 0)
    o = O.Oracle(d, 8, 1, O.DOT_FP64 | O.SVD_FP64)
    o.set_particles(x, v, m, material=mat)
    o.initialize()
    for _ in range(steps):
        o.advance_step(dt)
    p = o.particles()
    st = o.cg_stats()
    return dict(scene=np.array(scene), counts=np.int32(counts), dt=np.float64(dt), steps=np.int32(steps),
                x0=x, v0=v, m0=m, material0=mat, ids=p["ids"], x=p["x"], v=p["v"], F=p["F"], Jp=p["Jp"], volume=p["volume"],
                cg_iterations=np.int32(st["cg_iterations"]))


if __name__ == "__main__":
    out = os.path.dirname(os.path.abspath(__file__))
    for case in CASES:
        data = make(*case)
        path = os.path.join(out, case[0] + ".npz")
        np.savez_compressed(path, **data)
        print(path, len(data["ids"]), "particles")
