"""Golden fixtures (tests/golden/*.npz, written by make_fixtures.py from the
FP64 oracle): the oracle reproduces them exactly (CPU), the GPU path within
the FP32 tolerance (GPU)."""
import glob
import os

import numpy as np
import pytest

import non_ficks_diffusion_amd as M
import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURES = sorted(glob.glob(os.path.join(HERE, "golden", "*.npz")))


def load(path):
    with np.load(path, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a, float) - b) / max(np.linalg.norm(b), 1e-30))


@pytest.mark.parametrize("path", FIXTURES, ids=[os.path.basename(p) for p in FIXTURES])
def test_oracle_reproduces_fixture(path):
    f = load(path)
    d, x, v, m = M.scene_generate(str(f["scene"]), int(f["counts"]), 0)
    assert np.array_equal(x, f["x0"]) and np.array_equal(m, f["m0"])
    mat = f.get("material0")
    o = O.Oracle(d, 8, 1, O.DOT_FP64 | O.SVD_FP64)
    o.set_particles(x, v, m, material=mat)
    o.initialize()
    for _ in range(int(f["steps"])):
        o.advance_step(float(f["dt"]))
    p = o.particles()
    assert np.array_equal(p["ids"], f["ids"])
    for k in ("x", "v", "F", "Jp", "volume"):
        assert rel(p[k], f[k]) < 1e-12, k


@pytest.mark.gpu
@pytest.mark.parametrize("path", FIXTURES, ids=[os.path.basename(p) for p in FIXTURES])
def test_gpu_matches_fixture(path):
    f = load(path)
    d, x, v, m = M.scene_generate(str(f["scene"]), int(f["counts"]), 0)
    s = M.Scene(d, len(m))
    s.upload(x, v, m, material=f.get("material0"))
    s.initialize()
    for _ in range(int(f["steps"])):
        s.substep(float(f["dt"]))
    g = s.download()
    assert np.array_equal(g["ids"], f["ids"])
    for k in ("x", "v", "F", "Jp", "volume"):
        e = rel(g[k], f[k])
        print(os.path.basename(path), k, f"{e:.2e}")
        assert e < 1e-5, (k, e)
