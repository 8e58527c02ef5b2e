"""The allocation-free substep (mpm_substep_noalloc) against the handle path.

One substep of a scene on particle arrays that the caller holds in device
memory (here: buffers the test allocates through the HIP runtime the library
links), with every working buffer in the library's static scratch, must give
what mpm_create + mpm_upload_particles + mpm_initialize + mpm_substep give
for the same particles (the FP32 bar of one substep, 1e-5 relative L2: the
static scratch caps the block capacity at MPM_NOALLOC_MAX_BLOCKS, so the
node kernels' launch grid, and with it the grouping of the CG's FP64
partial sums, can differ from the handle's; both CG solves converge).  Two calls in a row, the second with
fewer particles and a smaller grid, so the second runs on the scratch the
first left behind; then the 2D path.  Sizes beyond the limits are refused
before any GPU work.
"""
import ctypes

import numpy as np
import pytest

import non_ficks_diffusion_amd as M

pytestmark = pytest.mark.gpu

E_INVALID, E_CAPACITY = -1, -3


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


def handle_path(d, x, v, m, dt):
    s = M.Scene(d, len(m))
    s.upload(x, v, m)
    s.initialize()
    st = s.substep(dt, stats=True)
    out = s.download()
    s.close()
    return out, st


class Hip:
    """The caller's side: device buffers and a stream through the HIP runtime
    the library itself links (libamdhip64.so.7), loaded with ctypes."""

    H2D, D2H = 1, 2

    def __init__(self):
        self.L = ctypes.CDLL("libamdhip64.so.7")
        self.L.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.L.hipFree.argtypes = [ctypes.c_void_p]
        self.L.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.L.hipStreamCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
        self.L.hipStreamSynchronize.argtypes = [ctypes.c_void_p]
        self.L.hipStreamDestroy.argtypes = [ctypes.c_void_p]
        self.L.hipDeviceSynchronize.argtypes = []

    def ok(self, rc):
        assert rc == 0, f"HIP error {rc}"

    def upload(self, a):
        a = np.ascontiguousarray(a)
        p = ctypes.c_void_p()
        self.ok(self.L.hipMalloc(ctypes.byref(p), max(a.nbytes, 4)))
        self.ok(self.L.hipMemcpy(p, a.ctypes.data, a.nbytes, self.H2D))
        return p.value

    def download(self, p, like):
        out = np.empty_like(like)
        self.ok(self.L.hipMemcpy(out.ctypes.data, p, out.nbytes, self.D2H))
        return out

    def free(self, p):
        self.ok(self.L.hipFree(p))


def noalloc_path(d, x, v, m, dt, hip, use_stream):
    n = len(m)
    host = {"x": np.ascontiguousarray(x, np.float32), "v": np.ascontiguousarray(v, np.float32),
            "F": np.tile(np.eye(d.dim, dtype=np.float32), (n, 1, 1)), "Jp": np.ones(n, np.float32),
            "mass": np.ascontiguousarray(m, np.float32), "volume": np.zeros(n, np.float32),
            "valid": np.full(n, 7, np.int32)}
    dev = {k: hip.upload(a) for k, a in host.items()}
    stream = None
    if use_stream:
        sp = ctypes.c_void_p()
        hip.ok(hip.L.hipStreamCreate(ctypes.byref(sp)))
        stream = sp.value
    st = M.substep_noalloc(d, n, dev["x"], dev["v"], dev["F"], dev["Jp"], dev["mass"], dev["volume"], dev["valid"],
                           dt, stream=stream)
    if stream is not None:
        hip.ok(hip.L.hipStreamSynchronize(stream))
        hip.ok(hip.L.hipStreamDestroy(stream))
    out = {k: hip.download(dev[k], host[k]) for k in host}
    for p in dev.values():
        hip.free(p)
    return out, st


def check(got, st_got, ref, st_ref, what):
    valid = got["valid"].astype(bool)
    assert set(np.unique(got["valid"])) <= {0, 1}, what
    assert np.array_equal(np.nonzero(valid)[0], ref["ids"]), what
    for k in ("x", "v", "F", "Jp", "volume"):
        e = rel(got[k][valid], ref[k])
        print(f"{what} {k}: {e:.2e}", flush=True)
        assert e < 1e-5, (what, k, e)
    assert st_got["live_particles"] == st_ref["live_particles"], what
    assert st_got["active_blocks"] == st_ref["active_blocks"], what
    # both solves converged; their iteration counts may differ by a few where
    # the residual sits at the FP32 noise floor near the tolerance
    for st in (st_got, st_ref):
        assert st["cg_converged"] == 1 and st["cg_residual"] <= 1e-6, (what, st)
    print(what, "cg iterations", st_got["cg_iterations"], "handle", st_ref["cg_iterations"], flush=True)


def test_noalloc_substep_matches_the_handle_path_twice():
    hip = Hip()
    calls = [("two_balls", 32, 0, 0, 1e-3, True), ("blob", 24, 1001, 2000, 2e-3, False)]
    sizes = []
    for name, counts, seed, n, dt, use_stream in calls:
        d, x, v, m = M.scene_generate(name, counts, seed, n)
        sizes.append(len(m))
        ref, st_ref = handle_path(d, x, v, m, dt)
        got, st_got = noalloc_path(d, x, v, m, dt, hip, use_stream)
        check(got, st_got, ref, st_ref, f"{name} {counts}^3 ({len(m)} particles)")
    assert sizes[1] < sizes[0]


def test_noalloc_substep_2d_matches_the_handle_path():
    """The 2D path of the entry point (8^2 blocks, 2x2 F) on the 2D disk."""
    hip = Hip()
    d, x, v, m = M.scene_generate("disk2d", 64, 0)
    assert d.dim == 2
    ref, st_ref = handle_path(d, x, v, m, 2e-3)
    got, st_got = noalloc_path(d, x, v, m, 2e-3, hip, True)
    check(got, st_got, ref, st_ref, f"disk2d 64^2 ({len(m)} particles)")


def test_noalloc_refuses_sizes_beyond_its_limits():
    d, x, v, m = M.scene_generate("two_balls", 32, 0)
    fake = 1  # never dereferenced: the sizes are refused before any GPU work
    with pytest.raises(M.MPMError) as e:
        M.substep_noalloc(d, M.NOALLOC_MAX_PARTICLES + 1, fake, fake, None, None, fake, fake, fake, 1e-3)
    assert e.value.code == E_CAPACITY
    big, _ = M.scene_preset("two_balls", M.NOALLOC_MAX_COUNTS + 4, 0)
    with pytest.raises(M.MPMError) as e:
        M.substep_noalloc(big, 100, fake, fake, None, None, fake, fake, fake, 1e-3)
    assert e.value.code == E_CAPACITY
    d.cg_max_iterations = 300
    with pytest.raises(M.MPMError) as e:
        M.substep_noalloc(d, 100, fake, fake, None, None, fake, fake, fake, 1e-3)
    assert e.value.code == E_INVALID
