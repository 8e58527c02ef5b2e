"""The C++ host driver (non_ficks_diffusion_amd/host): MPM_Driver's frame loop
over the C ABI, the reference's CLI options and frame files
(MPM_Driver.cpp:18-116, mpm_3d/main.cpp:10-32,
Read_Write/Read_Write_MPM_Particles.h:14-47)."""
import os
import subprocess

import numpy as np
import pytest

import non_ficks_diffusion_amd as M

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "non_ficks_diffusion_amd", "mpm_driver")

needs_driver = pytest.mark.skipif(not os.path.exists(DRIVER), reason="mpm_driver not built (run build())")


def read_frame(path, d):
    """Frame files: size_t count, then per particle X[d] V[d] mass volume
    (floats); `state` holds F[d*d], Jp (float), the material entry (int32)
    and the density (float) per particle."""
    raw = open(os.path.join(path, "particles"), "rb").read()
    n = int(np.frombuffer(raw[:8], np.uint64)[0])
    rec = np.frombuffer(raw[8:], np.float32).reshape(n, 2 * d + 2)
    st = open(os.path.join(path, "state"), "rb").read()
    assert int(np.frombuffer(st[:8], np.uint64)[0]) == n
    srec = np.frombuffer(st[8:], np.float32).reshape(n, d * d + 3)
    return {"x": rec[:, :d], "v": rec[:, d:2 * d], "mass": rec[:, 2 * d], "volume": rec[:, 2 * d + 1],
            "F": srec[:, :d * d].reshape(n, d, d), "Jp": srec[:, d * d],
            "material": srec[:, d * d + 1].copy().view(np.int32), "density": srec[:, d * d + 2]}


def run(args, timeout=300):
    return subprocess.run([DRIVER] + [str(a) for a in args], capture_output=True, text=True, timeout=timeout)


@needs_driver
def test_cli_usage_and_errors():
    r = run(["-h"])
    assert r.returncode == 0 and "-test_number" in r.stderr
    r = run(["-bogus"])
    assert r.returncode == 2 and "unknown option" in r.stderr
    r = run(["-d", "4"])
    assert r.returncode == 2
    r = run(["-levels", "2"])
    assert r.returncode == 2


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


@pytest.mark.gpu
@needs_driver
def test_driver_matches_library_frame(tmp_path):
    """One frame through the C++ MPM_Driver equals mpm_advance_to_time on
    the same scene (same kernels, same float dt sequence)."""
    out = tmp_path / "run"
    r = run(["-d", 3, "-test_number", 100, "-size", 32, "-last_frame", 1, "-o", out])
    assert r.returncode == 0, r.stderr
    f0 = read_frame(out / "0", 3)
    f1 = read_frame(out / "1", 3)
    desc, x, v, m = M.scene_generate("two_balls", 32, 0)
    assert f0["x"].shape[0] == len(m)
    np.testing.assert_array_equal(f0["x"], x)
    assert (f0["volume"] > 0).all(), "frame 0 must carry the volumes estimated at initialisation"
    s = M.Scene(desc, len(m))
    s.upload(x, v, m)
    s.initialize()
    target = float(np.float32(1.0) / np.float32(desc.frame_rate))
    t, n, _ = s.advance_to_time(0.0, target)
    g = s.download()
    assert g["x"].shape == f1["x"].shape
    for k in ("x", "v", "F", "Jp", "volume"):
        assert _rel(f1[k], g[k]) < 1e-5, k


@pytest.mark.gpu
@needs_driver
@pytest.mark.parametrize("dim,test,name,size", [(2, 16, "ref2d_test16", 64), (2, 19, "ref2d_test19", 64),
                                               (3, 19, "ref3d_test19", 32)])
def test_driver_material_scenes_match_library_frame(tmp_path, dim, test, name, size):
    """Reference tests 16 (two materials) and 19 (EOS fluid) through the C++
    MPM_Driver: the frame files carry each particle's material entry and, for
    an EOS particle, its density, equal to the library's state after the
    same frame."""
    out = tmp_path / "run"
    r = run(["-d", dim, "-test_number", test, "-size", size, "-last_frame", 1, "-o", out])
    assert r.returncode == 0, r.stderr
    f1 = read_frame(out / "1", dim)
    desc, x, v, m = M.scene_generate(name, size, 0)
    mat = M.scene_materials(name, size, 0)
    s = M.Scene(desc, len(m))
    s.upload(x, v, m, material=mat)
    s.initialize()
    s.advance_to_time(0.0, float(np.float32(1.0) / np.float32(desc.frame_rate)))
    g = s.download()
    s.close()
    assert np.array_equal(f1["material"], g["material"])
    eos = np.array([desc.materials()[k].eos for k in g["material"]], bool)
    for k in ("x", "v", "F"):
        assert _rel(f1[k], g[k]) < 1e-5, k
    if eos.any():
        assert _rel(f1["density"][eos], g["Jp"][eos]) < 1e-5
        assert np.all(f1["density"][eos] != 1.0)
    if (~eos).any():
        assert _rel(f1["Jp"][~eos], g["Jp"][~eos]) < 1e-5


@pytest.mark.gpu
@needs_driver
@pytest.mark.parametrize("test", [18, 16, 19])
def test_driver_staged_equals_fused(tmp_path, test):
    """-staged runs each MPM_Driver::Advance_Step stage as its own
    mpm_run_stage call; the result equals the fused substep to FP32 rounding
    (test 18 plastic balls, test 16 two materials, test 19 EOS fluid)."""
    a, b = tmp_path / "fused", tmp_path / "staged"
    r = run(["-d", 2, "-test_number", test, "-size", 64, "-last_frame", 1, "-o", a])
    assert r.returncode == 0, r.stderr
    r = run(["-d", 2, "-test_number", test, "-size", 64, "-last_frame", 1, "-o", b, "-staged"])
    assert r.returncode == 0, r.stderr
    fa, fb = read_frame(a / "1", 2), read_frame(b / "1", 2)
    assert fa["x"].shape == fb["x"].shape
    for k in ("x", "v", "F", "density"):
        assert _rel(fa[k], fb[k]) < 1e-4, k


@pytest.mark.gpu
@needs_driver
def test_driver_restart_is_lossless(tmp_path):
    """-restart 1 continues from frame 1's files (F and Jp included) and
    reaches frame 2 as the uninterrupted run does."""
    out = tmp_path / "run"
    r = run(["-d", 3, "-test_number", 18, "-size", 32, "-last_frame", 2, "-o", out])
    assert r.returncode == 0, r.stderr
    full = read_frame(out / "2", 3)
    os.rename(out / "2", tmp_path / "frame2_full")
    r = run(["-d", 3, "-test_number", 18, "-size", 32, "-last_frame", 2, "-o", out, "-restart", 1])
    assert r.returncode == 0, r.stderr
    again = read_frame(out / "2", 3)
    assert again["x"].shape == full["x"].shape
    for k in ("x", "v", "F", "Jp"):
        assert _rel(again[k], full[k]) < 1e-4, k


@pytest.mark.gpu
@needs_driver
def test_driver_writes_viewer_hierarchy(tmp_path):
    """Per frame, the viewer's hierarchy files (MPM_Example.cpp:1061-1062,
    MPM_Renderable.h:68-78, 125-168): one level of 4^3 blocks, each block's
    first cell index and its 64 flags, consistent with the dense flags of
    the same state."""
    out = tmp_path / "run"
    r = run(["-d", 3, "-test_number", 100, "-size", 32, "-last_frame", 1, "-o", out])
    assert r.returncode == 0, r.stderr
    head = np.frombuffer(open(out / "common" / "hierarchy.struct3d", "rb").read(), np.int32)
    assert list(head) == [1, 64, 4, 4, 4]
    grid = open(out / "common" / "fine_grid", "rb").read()
    assert list(np.frombuffer(grid[:12], np.int32)) == [32, 32, 32]
    assert list(np.frombuffer(grid[12:], np.float32)) == [0, 0, 0, 1, 1, 1]
    for frame in ("0", "1"):
        assert open(out / frame / "levels").read().strip() == "1"
        bo = open(out / frame / "block_offsets", "rb").read()
        nb = int(np.frombuffer(bo[:4], np.uint32)[0])
        idx = np.frombuffer(bo[4:], np.int32).reshape(nb, 3)
        fl = np.frombuffer(open(out / frame / "flags", "rb").read(), np.uint32).reshape(nb, 64)
        assert nb > 0 and (fl & 1).any()
        assert idx.min() >= 1 - 4 and idx.max() <= 32 + 1 and np.all((idx - 1) % 4 == 0)


@pytest.mark.gpu
@pytest.mark.parametrize("name,counts", [("two_balls", 32), ("disk2d", 64)])
def test_block_flags_match_dense_flags(name, counts):
    d, x, v, m = M.scene_generate(name, counts, 0)
    s = M.Scene(d, len(m))
    s.upload(x, v, m)
    s.initialize()
    s.substep(1e-3)
    origin, flags = s.read_block_flags()
    dense = s.read_flags()
    B = 4 if d.dim == 3 else 8
    rebuilt = np.zeros_like(dense)
    for b in range(len(origin)):
        for l in range(64):
            lc = [l >> 4, (l >> 2) & 3, l & 3] if d.dim == 3 else [l >> 3, l & 7]
            node = [origin[b][a] + lc[a] for a in range(d.dim)]
            if all(-1 <= i <= counts for i in node):
                rebuilt[tuple(i + 1 for i in node)] = flags[b, l]
    assert np.array_equal(rebuilt, dense)
    assert np.all(origin % B == 0)
    s.close()
