"""The fluid source (MPM_Example::Add_Fluid_Source, MPM_Example.cpp:977-1027):
EOS particles joining a running scene between substeps as waiting
particles, their volume from the grid in the next substep's
Process_Waiting_Particles.  Checked against the oracle (the same
particles added through its add_particles), launched against captured
substeps, from an empty start, and through the C++ driver's
-fluid_source run of the reference's 2D test 20.
"""
import os
import subprocess

import numpy as np
import pytest

import non_ficks_diffusion_amd as M
import oracle as O

pytestmark = pytest.mark.gpu

MATCHED = O.DOT_FP64 | O.SVD_FP64
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "non_ficks_diffusion_amd", "mpm_driver")


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


def emission(rng, k=10):
    """Ten particles of the 2D source: uniform in [0.45,0.55]^2, V=(0,-1),
    mass density 2 over 5e-4 each."""
    x = rng.uniform(0.45, 0.55, (k, 2)).astype(np.float32)
    v = np.tile(np.array([0.0, -1.0], np.float32), (k, 1))
    m = np.full(k, 2.0 * 5e-4, np.float32)
    return x, v, m


def test_emission_matches_oracle():
    d, _ = M.scene_preset("ref2d_test20", 64)
    rng = np.random.default_rng(20)
    x0, v0, m0 = emission(rng)
    steps, dt = 40, 2e-3
    s = M.Scene(d, 10 * (steps + 1))
    s.upload(x0, v0, m0)
    s.initialize()
    o = O.Oracle(d, 4, 4, MATCHED)
    o.set_particles(x0, v0, m0)
    o.initialize()
    for step in range(steps):
        xe, ve, me = emission(rng)
        s.add_particles(xe, ve, me)
        o.add_particles(xe, ve, me)
        s.substep(dt)
        o.advance_step(dt)
    g, r = s.download(), o.particles()
    assert len(g["ids"]) == 10 * (steps + 1) == len(r["ids"])
    assert np.array_equal(g["ids"], r["ids"])
    assert np.all(g["material"] == 0)
    errs = {k: rel(g[k], r[k]) for k in ("x", "v", "Jp", "volume")}
    print("fluid source, 40 substeps:", {k: f"{e:.1e}" for k, e in errs.items()})
    for k, e in errs.items():
        assert e < 1e-3, (k, e)
    # every particle has a volume, the densities moved off their initial 1
    assert np.all(g["volume"] > 0) and np.any(np.abs(g["Jp"] - 1.0) > 1e-4)
    s.close()


def test_graph_substeps_with_emission_match_launched():
    d, _ = M.scene_preset("ref2d_test20", 64)
    d.cg_max_iterations = 200  # captured substeps run a fixed cap
    runs = []
    for graph in (False, True):
        rng = np.random.default_rng(3)
        x0, v0, m0 = emission(rng)
        s = M.Scene(d, 400)
        s.upload(x0, v0, m0)
        s.initialize()
        s.use_graph(graph)
        for step in range(12):
            if step % 3 == 0:
                s.add_particles(*emission(rng))
            s.substep(2e-3)
        runs.append(s)
    a, b = runs[0].download(), runs[1].download()
    assert np.array_equal(a["ids"], b["ids"]) and len(a["ids"]) == 50
    # the same kernels on the same layout: bit for bit
    for k in ("x", "v", "Jp", "volume"):
        assert np.array_equal(b[k], a[k]), k
    for s in runs:
        s.close()


def test_empty_start_then_emission():
    d, x, v, m = M.scene_generate("ref2d_test20", 64)
    assert len(m) == 0
    s = M.Scene(d, 100)
    s.upload(x, v, m)
    s.initialize()
    rng = np.random.default_rng(1)
    for _ in range(3):
        s.add_particles(*emission(rng))
        s.substep(1e-3)
    g = s.download()
    assert list(g["ids"]) == list(range(30))
    assert np.all(g["volume"] > 0)
    assert s.stats()["live_particles"] == 30
    with pytest.raises(M.MPMError):
        s.add_particles(*emission(rng, 80))  # 30 + 80 > capacity 100
    s.close()


@pytest.mark.skipif(not os.path.exists(DRIVER), reason="mpm_driver not built (run build())")
def test_driver_fluid_source_run(tmp_path):
    """-fluid_source turns on the call the reference keeps commented out:
    test 20 emits ten particles per substep into its walled box."""
    out = tmp_path / "run"
    r = subprocess.run([DRIVER, "-d", "2", "-test_number", "20", "-size", "64", "-last_frame", "1", "-o", str(out),
                        "-fluid_source"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    substeps = int(r.stdout.split("substeps:")[1].split()[0])
    raw = open(out / "1" / "particles", "rb").read()
    n = int(np.frombuffer(raw[:8], np.uint64)[0])
    assert substeps > 0 and n == 10 * substeps
    rec = np.frombuffer(raw[8:], np.float32).reshape(n, 6)
    assert np.all(rec[:, 5] > 0)  # volumes
    # without the flag nothing is emitted (the reference's default)
    r = subprocess.run([DRIVER, "-d", "2", "-test_number", "20", "-size", "64", "-last_frame", "1", "-o",
                        str(tmp_path / "off")], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    assert int(np.frombuffer(open(tmp_path / "off" / "1" / "particles", "rb").read()[:8], np.uint64)[0]) == 0
