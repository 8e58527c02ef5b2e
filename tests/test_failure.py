"""Failure handling through the C ABI, every condition forced from the host
(no kernel is made to fault): non-finite or non-positive inputs are refused
with a status, a substep whose dt drives the state non-finite reports
MPM_E_NONFINITE and drops those particles (the reference drops them
silently, MPM_Example.cpp:650-660), particles outside the domain are
dropped at binning, a CG capped below convergence reports it in its stats
(the reference ignores the solver's result, MPM_Example.cpp:711), inverted
and singular F go through the SVD, the request server cancels queued
and running requests, and particles that need more active blocks than a
handle stores are refused at initialisation and make a substep a no-op
that reports MPM_E_CAPACITY.
"""
import numpy as np
import pytest

import non_ficks_diffusion_amd as M
import oracle as O

pytestmark = pytest.mark.gpu

E_INVALID, E_CAPACITY, E_NONFINITE = -1, -3, -5


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


def scene(name="two_balls", counts=24):
    d, x, v, m = M.scene_generate(name, counts, 0)
    return d, x.copy(), v.copy(), m.copy()


def test_non_finite_and_non_positive_inputs_are_refused():
    d, x, v, m = scene()
    s = M.Scene(d, len(m) + 10)
    for field, value in (("x", np.nan), ("v", np.inf), ("m", 0.0), ("m", -1.0), ("F", np.nan), ("Jp", np.inf),
                         ("Jp", 0.0)):
        xx, vv, mm = x.copy(), v.copy(), m.copy()
        F = np.tile(np.eye(3, dtype=np.float32), (len(m), 1, 1))
        Jp = np.ones(len(m), np.float32)
        {"x": xx[3], "v": vv[5], "m": mm[7:8], "F": F[2, 1], "Jp": Jp[4:5]}[field][...] = value
        with pytest.raises(M.MPMError) as e:
            s.upload(xx, vv, mm, F=F, Jp=Jp)
        assert e.value.code == E_INVALID, field
    s.upload(x, v, m)
    s.initialize()
    bad = x[:2].copy()
    bad[1, 0] = np.nan
    with pytest.raises(M.MPMError):
        s.add_particles(bad, v[:2], m[:2])
    s.substep(1e-3)  # the refused calls left the scene as it was
    assert s.stats()["live_particles"] == len(m)
    s.close()
    srv = M.Server(1 << 30, max_scenes=2, max_counts=64)
    vv = v.copy()
    vv[0, 1] = np.nan
    with pytest.raises(M.MPMError):
        srv.submit(d, x, vv, m, 1)
    srv.close()


def test_non_finite_state_is_reported_and_dropped():
    """dt = 1e30 drives v* and the CG to Inf/NaN: the substep's stats report
    MPM_E_NONFINITE, the non-finite particles are gone, and the handle keeps
    running on what is left."""
    d, x, v, m = scene()
    s = M.Scene(d, len(m))
    s.upload(x, v, m)
    s.initialize()
    with pytest.raises(M.MPMError) as e:
        s.substep(1e30, stats=True)
    assert e.value.code == E_NONFINITE
    g = s.download()
    assert np.all(np.isfinite(g["x"])) and len(g["ids"]) < len(m)
    s.close()


def test_particles_outside_the_domain_are_dropped():
    d, x, v, m = scene()
    x[0] = (1.2, 0.5, 0.5)
    x[1] = (-0.1, 0.5, 0.5)
    s = M.Scene(d, len(m))
    s.upload(x, v, m)
    s.initialize()
    s.substep(1e-3)
    g = s.download()
    assert 0 not in g["ids"] and 1 not in g["ids"] and len(g["ids"]) == len(m) - 2
    assert s.stats()["live_particles"] == len(m) - 2
    s.close()


def test_cg_capped_below_convergence_is_reported():
    d, x, v, m = scene()
    d.cg_max_iterations = 1
    d.cg_tolerance = 1e-12
    s = M.Scene(d, len(m))
    s.upload(x, v, m)
    s.initialize()
    st = s.substep(2e-3, stats=True)
    assert st["cg_converged"] == 0 and st["cg_iterations"] == 1
    d.cg_max_iterations = 50
    d.cg_tolerance = 1e-6
    s2 = M.Scene(d, len(m))
    s2.upload(x, v, m)
    s2.initialize()
    assert s2.substep(2e-3, stats=True)["cg_converged"] == 1
    s.close()
    s2.close()


def test_inverted_and_singular_deformation_gradients():
    """An inverted F (a reflection, det < 0) follows the FP64 oracle to the
    one-substep bar; a singular F (rank 1) gives a finite state.  At J = -1
    the lambda (J - 1) J term of P is large, so the material is softened and
    the CG given room to converge: the comparison is of converged solves."""
    d, x, v, m = scene()
    n = len(m)
    F = np.tile(np.eye(3, dtype=np.float32), (n, 1, 1))
    inv = np.arange(n) % 7 == 0
    F[inv] = np.diag([1.02, 0.97, -1.01]).astype(np.float32)  # det < 0
    d.material.plastic = 0
    d.material.mu0 *= 1e-3
    d.material.lambda0 *= 1e-3
    d.cg_tolerance = 1e-7
    d.cg_max_iterations = 200
    s = M.Scene(d, n)
    s.upload(x, v, m, F=F)
    s.initialize()
    st = s.substep(1e-3, stats=True)
    assert st["cg_converged"] == 1
    g = s.download()
    o = O.Oracle(d, 8, 4, O.DOT_FP64 | O.SVD_FP64)
    o.set_particles(x, v, m, F=F)
    o.initialize()
    o.advance_step(1e-3)
    r = o.particles()
    for k in ("x", "v", "F"):
        assert rel(g[k], r[k]) < 1e-5, (k, rel(g[k], r[k]))
    assert np.all(np.linalg.det(g["F"][inv]) < 0)
    s.close()
    Fs = np.tile(np.eye(3, dtype=np.float32), (n, 1, 1))
    Fs[::5] = np.outer([1.0, 0.5, 0.2], [0.3, 0.9, 0.1]).astype(np.float32)  # rank 1
    s = M.Scene(d, n)
    s.upload(x, v, m, F=Fs)
    s.initialize()
    s.substep(1e-3, stats=True)
    g = s.download()
    for k in ("x", "v", "F", "Jp"):
        assert np.all(np.isfinite(g[k])), k
    s.close()


def test_server_cancels_queued_and_running_requests():
    srv = M.Server(1 << 30, max_scenes=1, max_counts=64, keep_results=True)
    reqs = [M.scene_generate("two_balls", 32, 0, 3000), M.scene_generate("blob", 32, 1003, 2000),
            M.scene_generate("slab", 32, 1004, 2500)]
    ids = [srv.submit(d, x, v, m, 2) for d, x, v, m in reqs]
    srv.tick()
    assert srv.poll(ids[0])["state"] == M.REQ_RUNNING and srv.poll(ids[1])["state"] == M.REQ_QUEUED
    srv.cancel(ids[1])  # queued
    assert srv.poll(ids[1])["state"] == M.REQ_CANCELLED
    srv.cancel(ids[0])  # running: leaves the arena, the entry goes to the next request
    assert srv.poll(ids[0])["state"] == M.REQ_CANCELLED
    t = srv.tick()
    assert t["admitted"] == 1 and srv.poll(ids[2])["state"] == M.REQ_RUNNING
    while srv.poll(ids[2])["state"] != M.REQ_DONE:
        srv.tick()
    # the survivor's result is that of the scene alone
    d, x, v, m = reqs[2]
    s = M.Scene(d, len(m))
    s.upload(x, v, m)
    s.initialize()
    t = 0.0
    for f in range(2):
        t, _, _ = s.advance_to_time(t, float(np.float32(f + 1) / np.float32(24.0)))
    assert rel(srv.download(ids[2])["x"], s.download()["x"]) < 1e-4
    s.close()
    with pytest.raises(M.MPMError):
        srv.cancel(ids[2])  # already finished
    for rid in ids:
        srv.release(rid)
    srv.close()


def block_lattice(z_layers):
    """One particle inside each 4^3 block of a 128^3 grid, blocks 1..30 in x
    and y and 1..z_layers in z: every particle's stencil stays in its block,
    so the lattice needs exactly 900 * z_layers active blocks."""
    g = np.stack(np.meshgrid(np.arange(1, 31), np.arange(1, 31), np.arange(1, z_layers + 1), indexing="ij"),
                 -1).reshape(-1, 3)
    x = ((4 * g + 2.3) / 128).astype(np.float32)
    return x, np.zeros_like(x), np.full(len(x), 1e-3, np.float32)


def test_block_capacity_overflow_is_refused_and_changes_nothing():
    d, _ = M.scene_preset("two_balls", 128, 0)
    x, v, m = block_lattice(6)  # 5400 blocks; a 5400-particle handle stores 4096
    s = M.Scene(d, len(m))
    s.upload(x, v, m)
    with pytest.raises(M.MPMError) as e:
        s.initialize()
    assert e.value.code == E_CAPACITY
    with pytest.raises(M.MPMError) as e:
        s.substep(1e-4)  # not initialised
    # the handle takes a set that fits, and gives what a fresh handle gives
    x1, v1, m1 = block_lattice(1)
    s.upload(x1, v1, m1)
    s.initialize()
    s.substep(1e-4)
    got = s.download()
    ref_s = M.Scene(d, len(m1))
    ref_s.upload(x1, v1, m1)
    ref_s.initialize()
    ref_s.substep(1e-4)
    ref = ref_s.download()
    ref_s.close()
    assert np.array_equal(got["ids"], ref["ids"])
    for f in ("x", "v", "F"):
        assert rel(got[f], ref[f]) < 1e-6, f
    # particles that join later push the blocks over capacity: the substep is
    # a no-op that reports it, and the particles already there keep their state
    x2, v2, m2 = block_lattice(6)
    extra = x2[:, 2] > 8 / 128  # blocks z = 2..6: 4500 more
    before = s.download()
    s.add_particles(x2[extra], v2[extra], m2[extra])
    with pytest.raises(M.MPMError) as e:
        s.substep(1e-4)
    assert e.value.code == E_CAPACITY
    after = s.download()
    keep = np.isin(after["ids"], before["ids"])
    assert keep.sum() == len(before["ids"])
    for f in ("x", "v", "F", "Jp", "volume"):
        assert np.array_equal(after[f][keep], before[f]), f
    s.close()


@pytest.mark.parametrize("graph", [False, True])
def test_overflow_after_capped_solves_is_reported(graph):
    """A CG solve that runs to its cap (here: one iteration) leaves its last
    done poll unread until the rest of the substep is queued; the overflow
    flag that poll carries must still make the substep report
    MPM_E_CAPACITY.  A lattice
    of 3,600 particles, one per 4^3 block (3,600 active blocks of the 4,096
    the handle stores), moves one cell diagonally in a substep, after which
    every particle's stencil spans 8 blocks (4,805 in all).  Replayed
    substeps read no poll: the next call that reads the device state reports
    it (the max speed, the device frame loop)."""
    d, _ = M.scene_preset("two_balls", 128, 0)
    d.cg_max_iterations = 1  # every solve is one chunk: its poll is left for after G2P
    d.cg_tolerance = 1e-12
    d.nplanes = 0
    d.gravity[:] = (0.0, 0.0, 0.0)
    x, _, m = block_lattice(4)
    dt = 1e-4
    cell_per_dt = np.float32(1.0 / 128 / dt)
    rng = np.random.default_rng(8)
    v = (cell_per_dt + 0.2 * rng.standard_normal(x.shape)).astype(np.float32)
    s = M.Scene(d, len(m))
    s.upload(x, v, m)
    s.initialize()
    if graph:
        s.use_graph(True)
    s.substep(dt)  # the particles cross into the next cell
    assert s.stats()["cg_iterations"] <= 1
    if graph:
        s.substep(dt)  # a replay: no poll
    else:
        with pytest.raises(M.MPMError) as e:
            s.substep(dt)
        assert e.value.code == E_CAPACITY
    with pytest.raises(M.MPMError) as e:
        s.max_particle_velocity()
    assert e.value.code == E_CAPACITY
    if graph:
        with pytest.raises(M.MPMError) as e:
            s.advance_to_time(0.0, 1e-3)
        assert e.value.code == E_CAPACITY
    s.close()


def test_call_order_and_argument_errors():
    # every condition is checked on the host and answered with a status; the
    # handle stays usable afterwards
    E_NOT_READY = -4
    d, x, v, m = scene()
    n = len(m)
    s = M.Scene(d, n)

    def code(f, *a):
        with pytest.raises(M.MPMError) as e:
            f(*a)
        return e.value.code

    assert code(s.initialize) == E_NOT_READY  # nothing uploaded
    big = np.concatenate([x, x[:1]]), np.concatenate([v, v[:1]]), np.concatenate([m, m[:1]])
    assert code(s.upload, *big) == E_CAPACITY  # one more than the capacity
    s.upload(x, v, m)
    assert code(s.substep, 1e-4) == E_NOT_READY  # before initialize
    assert code(s.time_cg_operator, 5) == E_NOT_READY  # before a substep
    s.initialize()
    for dt in (0.0, -1e-4, float("nan"), float("inf")):
        assert code(s.substep, dt) == E_INVALID
    s.substep(1e-4)
    assert code(s.download, n - 1) == E_CAPACITY  # cap below the live count
    got = s.download()
    assert len(got["ids"]) == n
    # the CG operator reads the records of the constitutive update on the
    # current binning: a staged CG needs CONSTITUTIVE after the last BIN, and
    # operator timing needs records that stage timing has not rewritten
    s.run_stage(M.STAGE_BIN)
    s.run_stage(M.STAGE_P2G)
    assert code(s.run_stage, M.STAGE_CG, 1e-4) == E_NOT_READY
    s.run_stage(M.STAGE_CONSTITUTIVE)
    s.run_stage(M.STAGE_EXPLICIT_FORCE, 1e-4)
    s.run_stage(M.STAGE_COLLISION, 1e-4)
    s.run_stage(M.STAGE_CG, 1e-4)
    s.run_stage(M.STAGE_G2P, 1e-4)
    assert s.time_cg_operator(2) > 0.0
    assert s.time_stage(M.STAGE_CONSTITUTIVE, 2) > 0.0
    assert code(s.time_cg_operator, 2) == E_NOT_READY
    s.substep(1e-4)
    assert s.time_cg_operator(2) > 0.0
    s.close()

    b = M.Batch([d, d], [n, n])
    b.upload(0, x, v, m)
    b.upload(1, x, v, m)
    b.initialize()
    assert code(b.upload, 0, x, v, m) == E_NOT_READY  # the batch is built
    b.substep(1e-4)
    assert code(b.stats, 2) == E_INVALID  # no such scene
    b.close()


def test_scene_emptied_by_advection_keeps_running_and_refills():
    # every particle leaves the domain in the first substep (fast and near the
    # x = 1 edge); later substeps run on an empty scene, and particles added
    # afterwards are simulated as in a scene that started with them
    d, _ = M.scene_preset("two_balls", 32, 0)
    rng = np.random.default_rng(3)
    n = 200
    x = np.column_stack([rng.uniform(0.93, 0.95, n), rng.uniform(0.4, 0.6, n), rng.uniform(0.4, 0.6, n)])
    x = x.astype(np.float32)
    v = np.tile(np.float32([400.0, 0.0, 0.0]), (n, 1))
    m = np.full(n, 1e-3, np.float32)
    s = M.Scene(d, 2 * n)
    s.upload(x, v, m)
    s.initialize()
    s.substep(1e-3)  # 0.4 in x: everyone leaves
    assert len(s.download()["ids"]) == 0
    for _ in range(3):
        st = s.substep(1e-3, stats=True)
        assert st["live_particles"] == 0 and st["active_blocks"] == 0
    assert s.max_particle_velocity() == 0.0
    x2 = x - np.float32([0.5, 0.0, 0.0])
    v2 = np.zeros_like(v)
    s.add_particles(x2, v2, m)
    for _ in range(3):
        s.substep(1e-4)
    got = s.download()
    ref_s = M.Scene(d, n)
    ref_s.upload(x2, v2, m)
    ref_s.initialize()
    for _ in range(3):
        ref_s.substep(1e-4)
    ref = ref_s.download()
    ref_s.close()
    assert len(got["ids"]) == n and np.array_equal(got["ids"] - n, ref["ids"])
    for f in ("x", "v", "F", "volume"):
        assert rel(got[f], ref[f]) < 1e-5, (f, rel(got[f], ref[f]))
    s.close()
