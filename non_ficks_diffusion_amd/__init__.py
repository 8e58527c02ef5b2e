"""Python mirror of the libmpm_gpu C ABI (include/mpm_gpu.h).

The MI355X MPM snow substep is C++/HIP behind a C ABI; this module binds
that ABI with ctypes for tests and benchmarks.  It mirrors the reference's
MPM_Example / MPM_Driver surface (MPM_Example.h:229-336,
MPM_Driver.h:13-36): a `Scene` owns the device-resident particles and
grid of one simulation; `initialize()` is MPM_Driver::Initialize,
`substep(dt)` is MPM_Driver::Advance_Step, `run_stage` runs one of its
stages.  There is no CPU fallback: importing works anywhere, but every
call that reaches the device raises if the HIP library is missing or the
runtime has no GPU.
"""
from __future__ import annotations

import ctypes
import os
from typing import Dict, Optional, Tuple

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmpm_gpu.so")

MPM_MAX_PLANES = 8
MPM_MAX_MATERIALS = 8
MPM_FLAG_INTERIOR = 1

STAGE_BIN, STAGE_P2G, STAGE_WAITING, STAGE_CONSTITUTIVE = 0, 1, 2, 3
STAGE_EXPLICIT_FORCE, STAGE_COLLISION, STAGE_CG, STAGE_G2P = 4, 5, 6, 7

CH_MASS, CH_VELOCITY, CH_VELOCITY_STAR, CH_FORCE = 0, 1, 2, 3
CH_CG_R, CH_CG_S, CH_CG_Q = 4, 5, 6


class MPMError(RuntimeError):
    """A libmpm_gpu call returned a non-zero status."""

    def __init__(self, code: int, msg: str):
        super().__init__(f"libmpm_gpu error {code}: {msg}")
        self.code = code


class Plane(ctypes.Structure):
    _fields_ = [("normal", ctypes.c_float * 3), ("surface", ctypes.c_float * 3),
                ("mu", ctypes.c_float), ("sticky", ctypes.c_int32)]


class Material(ctypes.Structure):
    _fields_ = [("mu0", ctypes.c_float), ("lambda0", ctypes.c_float),
                ("stretching_yield", ctypes.c_float), ("compression_yield", ctypes.c_float),
                ("hardening_factor", ctypes.c_float), ("hardening_max", ctypes.c_float),
                ("plastic", ctypes.c_int32), ("eos", ctypes.c_int32), ("bulk_modulus", ctypes.c_float),
                ("gamma", ctypes.c_float), ("density", ctypes.c_float)]


class SceneDesc(ctypes.Structure):
    _fields_ = [("dim", ctypes.c_int32), ("counts", ctypes.c_int32),
                ("gravity", ctypes.c_float * 3), ("flip", ctypes.c_float), ("cfl", ctypes.c_float),
                ("min_dt", ctypes.c_float), ("max_dt", ctypes.c_float), ("frame_rate", ctypes.c_float),
                ("cg_tolerance", ctypes.c_float), ("cg_max_iterations", ctypes.c_int32),
                ("cg_min_iterations", ctypes.c_int32), ("nplanes", ctypes.c_int32),
                ("planes", Plane * MPM_MAX_PLANES), ("material", Material),
                ("crop_ghost_x", ctypes.c_int32), ("project_mode", ctypes.c_int32),
                ("nmaterials", ctypes.c_int32), ("extra_materials", Material * (MPM_MAX_MATERIALS - 1))]

    def materials(self):
        """The scene's material table (entry 0 is `material`)."""
        return [self.material] + [self.extra_materials[k] for k in range(max(1, self.nmaterials) - 1)]

    def copy(self) -> "SceneDesc":
        d = SceneDesc()
        ctypes.memmove(ctypes.byref(d), ctypes.byref(self), ctypes.sizeof(SceneDesc))
        return d


class StepStats(ctypes.Structure):
    _fields_ = [("cg_iterations", ctypes.c_int32), ("cg_multiplies", ctypes.c_int32),
                ("cg_converged", ctypes.c_int32), ("active_blocks", ctypes.c_int32),
                ("live_particles", ctypes.c_int64), ("cg_residual", ctypes.c_double),
                ("max_speed", ctypes.c_double)]

    def as_dict(self) -> Dict[str, float]:
        return {k: getattr(self, k) for k, _ in self._fields_}


class LaunchConfig(ctypes.Structure):
    _fields_ = [("block_groups", ctypes.c_int32), ("node_groups", ctypes.c_int32),
                ("particle_groups", ctypes.c_int32), ("operator_mode", ctypes.c_int32),
                ("operator_groups", ctypes.c_int32), ("from_cache", ctypes.c_int32),
                ("block_ms", ctypes.c_float), ("node_ms", ctypes.c_float), ("particle_ms", ctypes.c_float),
                ("operator_ms", ctypes.c_float)]

    def as_dict(self) -> Dict[str, float]:
        return {k: getattr(self, k) for k, _ in self._fields_}


MPM_TUNE_CACHED = 1


class ServeConfig(ctypes.Structure):
    _fields_ = [("dim", ctypes.c_int32), ("max_scenes", ctypes.c_int32), ("max_counts", ctypes.c_int32),
                ("cg_max_iterations", ctypes.c_int32), ("hbm_budget_bytes", ctypes.c_int64),
                ("graph", ctypes.c_int32), ("keep_results", ctypes.c_int32),
                ("max_tick_substeps", ctypes.c_int32), ("tune", ctypes.c_int32)]


class RequestStatus(ctypes.Structure):
    _fields_ = [("state", ctypes.c_int32), ("frames_done", ctypes.c_int32), ("frames_total", ctypes.c_int32),
                ("substeps", ctypes.c_int32), ("particles", ctypes.c_int64), ("bytes", ctypes.c_int64),
                ("sim_time", ctypes.c_double), ("submitted", ctypes.c_double), ("admitted", ctypes.c_double),
                ("finished", ctypes.c_double)]

    def as_dict(self) -> Dict[str, float]:
        return {k: getattr(self, k) for k, _ in self._fields_}


class TickStats(ctypes.Structure):
    _fields_ = [("admitted", ctypes.c_int32), ("finished", ctypes.c_int32), ("resident", ctypes.c_int32),
                ("queued", ctypes.c_int32), ("substeps", ctypes.c_int32), ("frames", ctypes.c_int32),
                ("particle_substeps", ctypes.c_int64), ("bytes_in_use", ctypes.c_int64), ("ms", ctypes.c_double)]

    def as_dict(self) -> Dict[str, float]:
        return {k: getattr(self, k) for k, _ in self._fields_}


REQ_QUEUED, REQ_RUNNING, REQ_DONE, REQ_CANCELLED = 0, 1, 2, 3

_FP = ctypes.POINTER(ctypes.c_float)
_IP = ctypes.POINTER(ctypes.c_int32)
_UP = ctypes.POINTER(ctypes.c_uint32)
_LP = ctypes.POINTER(ctypes.c_int64)
_H = ctypes.c_void_p

# symbol -> (restype, argtypes); every function declared in include/mpm_gpu.h
SIGNATURES = {
    "mpm_last_error": (ctypes.c_char_p, []),
    "mpm_device_info": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int32]),
    "mpm_create": (ctypes.c_int, [ctypes.POINTER(SceneDesc), ctypes.c_int64, ctypes.POINTER(_H)]),
    "mpm_destroy": (ctypes.c_int, [_H]),
    "mpm_upload_particles": (ctypes.c_int, [_H, ctypes.c_int64, _FP, _FP, _FP, _FP, _FP, _FP]),
    "mpm_download_particles": (ctypes.c_int, [_H, ctypes.c_int64, _FP, _FP, _FP, _FP, _FP, _FP, _IP, _LP]),
    "mpm_add_particles": (ctypes.c_int, [_H, ctypes.c_int32, ctypes.c_int64, _FP, _FP, _FP, _IP]),
    "mpm_upload_particles_ex": (ctypes.c_int, [_H, ctypes.c_int64, _FP, _FP, _FP, _FP, _FP, _FP, _IP]),
    "mpm_download_particles_ex": (ctypes.c_int, [_H, ctypes.c_int64, _FP, _FP, _FP, _FP, _FP, _FP, _IP, _IP, _LP]),
    "mpm_initialize": (ctypes.c_int, [_H]),
    "mpm_substep": (ctypes.c_int, [_H, ctypes.c_float, ctypes.POINTER(StepStats)]),
    "mpm_advance_to_time": (ctypes.c_int, [_H, ctypes.c_double, ctypes.c_double, _IP,
                                           ctypes.POINTER(ctypes.c_double), ctypes.POINTER(StepStats)]),
    "mpm_max_particle_velocity": (ctypes.c_int, [_H, _FP]),
    "mpm_synchronize": (ctypes.c_int, [_H]),
    "mpm_run_stage": (ctypes.c_int, [_H, ctypes.c_int32, ctypes.c_float]),
    "mpm_read_grid": (ctypes.c_int, [_H, ctypes.c_int32, _FP, ctypes.c_int64]),
    "mpm_read_flags": (ctypes.c_int, [_H, _UP, ctypes.c_int64]),
    "mpm_read_block_flags": (ctypes.c_int, [_H, ctypes.c_int64, _IP, _UP, _IP]),
    "mpm_read_bins": (ctypes.c_int, [_H, ctypes.c_int64, _IP, _IP, ctypes.c_int64, _IP, _IP, _IP, _LP]),
    "mpm_cg_stats": (ctypes.c_int, [_H, ctypes.POINTER(StepStats)]),
    "mpm_enable_timing": (ctypes.c_int, [_H, ctypes.c_int32]),
    "mpm_kernel_times": (ctypes.c_int, [_H, ctypes.c_int32, ctypes.POINTER(ctypes.c_double), _LP]),
    "mpm_kernel_class_name": (ctypes.c_char_p, [ctypes.c_int32]),
    "mpm_kernel_class_count": (ctypes.c_int32, []),
    "mpm_time_cg_operator": (ctypes.c_int, [_H, ctypes.c_int32, ctypes.POINTER(ctypes.c_float)]),
    "mpm_measure_copy_bandwidth": (ctypes.c_int, [ctypes.c_int64, ctypes.c_int32, ctypes.POINTER(ctypes.c_double)]),
    "mpm_time_stage": (ctypes.c_int, [_H, ctypes.c_int32, ctypes.c_int32, ctypes.POINTER(ctypes.c_float)]),
    "mpm_get_launch_config": (ctypes.c_int, [_H, ctypes.POINTER(LaunchConfig)]),
    "mpm_set_launch_config": (ctypes.c_int, [_H, ctypes.POINTER(LaunchConfig)]),
    "mpm_autotune": (ctypes.c_int, [_H, _FP, ctypes.c_int32, ctypes.POINTER(LaunchConfig)]),
    "mpm_set_graph": (ctypes.c_int, [_H, ctypes.c_int32]),
    "mpm_serve_create": (ctypes.c_int, [ctypes.POINTER(ServeConfig), ctypes.POINTER(_H)]),
    "mpm_serve_destroy": (ctypes.c_int, [_H]),
    "mpm_serve_arena": (ctypes.c_int, [_H, _LP, _LP, _LP, _LP]),
    "mpm_serve_submit": (ctypes.c_int, [_H, ctypes.POINTER(SceneDesc), ctypes.c_int64, _FP, _FP, _FP, ctypes.c_int32,
                                        _LP]),
    "mpm_serve_submit_ex": (ctypes.c_int, [_H, ctypes.POINTER(SceneDesc), ctypes.c_int64, _FP, _FP, _FP, _IP,
                                           ctypes.c_int32, _LP]),
    "mpm_serve_tick": (ctypes.c_int, [_H, ctypes.POINTER(TickStats)]),
    "mpm_serve_poll": (ctypes.c_int, [_H, ctypes.c_int64, ctypes.POINTER(RequestStatus)]),
    "mpm_serve_download": (ctypes.c_int, [_H, ctypes.c_int64, ctypes.c_int64, _FP, _FP, _FP, _FP, _FP, _FP, _IP,
                                          _LP]),
    "mpm_serve_release": (ctypes.c_int, [_H, ctypes.c_int64]),
    "mpm_serve_cancel": (ctypes.c_int, [_H, ctypes.c_int64]),
    "mpm_serve_frame_times": (ctypes.c_int, [_H, ctypes.c_int64, ctypes.c_int32, ctypes.POINTER(ctypes.c_double),
                                             _IP]),
    "mpm_batch_create": (ctypes.c_int, [ctypes.POINTER(SceneDesc), _LP, ctypes.c_int32, ctypes.POINTER(_H)]),
    "mpm_batch_size": (ctypes.c_int32, [_H]),
    "mpm_batch_upload": (ctypes.c_int, [_H, ctypes.c_int32, ctypes.c_int64, _FP, _FP, _FP, _FP, _FP, _FP]),
    "mpm_batch_upload_ex": (ctypes.c_int, [_H, ctypes.c_int32, ctypes.c_int64, _FP, _FP, _FP, _FP, _FP, _FP, _IP]),
    "mpm_batch_substep": (ctypes.c_int, [_H, _FP]),
    "mpm_batch_max_velocity": (ctypes.c_int, [_H, _FP]),
    "mpm_batch_stats": (ctypes.c_int, [_H, ctypes.c_int32, ctypes.POINTER(StepStats)]),
    "mpm_batch_download": (ctypes.c_int, [_H, ctypes.c_int32, ctypes.c_int64, _FP, _FP, _FP, _FP, _FP, _FP, _IP, _LP]),
    "mpm_batch_download_ex": (ctypes.c_int, [_H, ctypes.c_int32, ctypes.c_int64, _FP, _FP, _FP, _FP, _FP, _FP, _IP,
                                             _IP, _LP]),
    "mpm_batch_advance_to_time": (ctypes.c_int, [_H, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double),
                                                 ctypes.POINTER(ctypes.c_double), _IP]),
    "mpm_substep_noalloc": (ctypes.c_int, [ctypes.POINTER(SceneDesc), ctypes.c_int64] + [ctypes.c_void_p] * 7
                            + [ctypes.c_float, ctypes.c_void_p, ctypes.POINTER(StepStats)]),
    "mpm_scene_preset": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int32, ctypes.c_uint64,
                                        ctypes.POINTER(SceneDesc), _LP]),
    "mpm_scene_generate": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int32, ctypes.c_uint64, ctypes.c_int64,
                                          _FP, _FP, _FP, _LP]),
    "mpm_scene_preset_sized": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int32, ctypes.c_uint64, ctypes.c_int64,
                                              ctypes.POINTER(SceneDesc), _LP]),
    "mpm_scene_generate_sized": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int32, ctypes.c_uint64, ctypes.c_int64,
                                                ctypes.c_int64, _FP, _FP, _FP, _LP]),
    "mpm_scene_materials": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int32, ctypes.c_uint64, ctypes.c_int64,
                                           ctypes.c_int64, _IP, _LP]),
}

_lib: Optional[ctypes.CDLL] = None


def lib() -> ctypes.CDLL:
    """Loads libmpm_gpu.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with __graft_entry__.build()")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _check(rc: int) -> None:
    if rc != 0:
        raise MPMError(rc, lib().mpm_last_error().decode())


def _fp(a: Optional[np.ndarray]):
    return None if a is None else a.ctypes.data_as(_FP)


def _f32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float32)


def _i32(a) -> Optional[np.ndarray]:
    return None if a is None else np.ascontiguousarray(a, dtype=np.int32)


def _ip(a: Optional[np.ndarray]):
    return None if a is None else a.ctypes.data_as(_IP)


def _particle_arrays(cap: int, d: int) -> Dict[str, np.ndarray]:
    return {"x": np.zeros((cap, d), np.float32), "v": np.zeros((cap, d), np.float32),
            "F": np.zeros((cap, d, d), np.float32), "Jp": np.zeros(cap, np.float32),
            "mass": np.zeros(cap, np.float32), "volume": np.zeros(cap, np.float32),
            "ids": np.zeros(cap, np.int32), "material": np.zeros(cap, np.int32)}


def device_info() -> str:
    buf = ctypes.create_string_buffer(256)
    _check(lib().mpm_device_info(buf, 256))
    return buf.value.decode()


def scene_preset(name: str, counts: int = 0, seed: int = 0, n: int = 0) -> Tuple[SceneDesc, int]:
    """Scene description and particle count of a named scene; n > 0 sizes
    the scene to exactly n particles (mpm_scene_preset_sized)."""
    d = SceneDesc()
    out = ctypes.c_int64()
    _check(lib().mpm_scene_preset_sized(name.encode(), counts, seed, n, ctypes.byref(d), ctypes.byref(out)))
    return d, out.value


def scene_generate(name: str, counts: int = 0, seed: int = 0, n: int = 0):
    """Returns (desc, x[n,dim], v[n,dim], mass[n]) of a named scene (scene hook
    Initialize_Particles, mpm_3d/Standard_Tests/Standard_Tests.h:47-154);
    n > 0 sizes the scene to exactly n particles."""
    d, count = scene_preset(name, counts, seed, n)
    x = np.zeros((count, d.dim), np.float32)
    v = np.zeros((count, d.dim), np.float32)
    m = np.zeros(count, np.float32)
    out = ctypes.c_int64()
    _check(lib().mpm_scene_generate_sized(name.encode(), counts, seed, n, count, _fp(x), _fp(v), _fp(m),
                                          ctypes.byref(out)))
    return d, x, v, m


def scene_materials(name: str, counts: int = 0, seed: int = 0, n: int = 0) -> np.ndarray:
    """Material table entry of every particle of a named scene, in the order
    scene_generate returns them (all 0 but in scenes that mix materials)."""
    _, count = scene_preset(name, counts, seed, n)
    out = np.zeros(count, np.int32)
    k = ctypes.c_int64()
    _check(lib().mpm_scene_materials(name.encode(), counts, seed, n, count, _ip(out), ctypes.byref(k)))
    return out


NOALLOC_MAX_PARTICLES, NOALLOC_MAX_COUNTS, NOALLOC_MAX_BLOCKS = 65536, 64, 2048


def substep_noalloc(desc: SceneDesc, n: int, x: int, v: int, F: Optional[int], Jp: Optional[int], mass: int,
                    volume: int, valid: int, dt: float, stream: Optional[int] = None, stats: bool = True):
    """One substep of one scene on particle arrays in device memory that the
    caller owns (device addresses), allocating nothing (mpm_substep_noalloc):
    x, v, F, Jp and volume are overwritten with the state after the substep,
    valid[i] = 1 for particles still inside the domain.  stream: a hipStream_t
    address or None.  Returns the step stats (stats=True waits for the
    stream) or None."""
    d = desc.copy()
    st = StepStats() if stats else None
    _check(lib().mpm_substep_noalloc(ctypes.byref(d), int(n), x, v, F, Jp, mass, volume, valid, float(dt), stream,
                                     ctypes.byref(st) if st is not None else None))
    return st.as_dict() if st is not None else None


def kernel_class_names():
    L = lib()
    return [L.mpm_kernel_class_name(k).decode() for k in range(L.mpm_kernel_class_count())]


def measure_copy_bandwidth(nbytes: int = 1 << 30, reps: int = 10) -> float:
    """GB/s of a streaming device-to-device copy of nbytes (read plus write),
    the measured bandwidth roof of the current device."""
    gbs = ctypes.c_double()
    _check(lib().mpm_measure_copy_bandwidth(int(nbytes), int(reps), ctypes.byref(gbs)))
    return gbs.value


class _LaunchControl:
    """Launch shapes, autotuner and captured substeps of a handle."""

    def launch_config(self) -> Dict[str, float]:
        g = LaunchConfig()
        _check(lib().mpm_get_launch_config(self._h, ctypes.byref(g)))
        return g.as_dict()

    def set_launch_config(self, block_groups: int, node_groups: int, particle_groups: int,
                          operator_mode: Optional[int] = None, operator_groups: Optional[int] = None) -> None:
        cur = self.launch_config()
        g = LaunchConfig(block_groups, node_groups, particle_groups,
                         cur["operator_mode"] if operator_mode is None else operator_mode,
                         cur["operator_groups"] if operator_groups is None else operator_groups, 0,
                         0.0, 0.0, 0.0, 0.0)
        _check(lib().mpm_set_launch_config(self._h, ctypes.byref(g)))

    def autotune(self, dt, cached: bool = False) -> Dict[str, float]:
        """Times candidate launch shapes on the current particles (state is
        not advanced) and keeps the fastest; cached: reuse an earlier result
        of the same size bucket."""
        d = np.ascontiguousarray(np.broadcast_to(np.asarray(dt, np.float32), (self._nscenes(),)), dtype=np.float32)
        g = LaunchConfig()
        _check(lib().mpm_autotune(self._h, _fp(d), MPM_TUNE_CACHED if cached else 0, ctypes.byref(g)))
        return g.as_dict()

    def use_graph(self, on: bool = True) -> None:
        """Each substep replays a captured HIP graph of its launch sequence."""
        _check(lib().mpm_set_graph(self._h, 1 if on else 0))


class Scene(_LaunchControl):
    """One device-resident simulation (MPM_Example + MPM_Driver on the GPU)."""

    def _nscenes(self) -> int:
        return 1

    def __init__(self, desc: SceneDesc, capacity: int):
        self.desc = desc.copy()
        self.dim = desc.dim
        self.counts = desc.counts
        self.capacity = int(capacity)
        h = _H()
        _check(lib().mpm_create(ctypes.byref(self.desc), int(capacity), ctypes.byref(h)))
        self._h = h

    def close(self) -> None:
        if getattr(self, "_h", None):
            lib().mpm_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # -- particles ---------------------------------------------------------
    def upload(self, x, v, mass, F=None, Jp=None, volume=None, material=None) -> None:
        """material: entry of the scene's material table per particle (None:
        all 0); an EOS particle's Jp is its density (None: its material's)."""
        x, v, mass = _f32(x), _f32(v), _f32(mass)
        n = mass.shape[0]
        F = None if F is None else _f32(F)
        Jp = None if Jp is None else _f32(Jp)
        volume = None if volume is None else _f32(volume)
        material = _i32(material)
        _check(lib().mpm_upload_particles_ex(self._h, n, _fp(x), _fp(v), _fp(F), _fp(Jp), _fp(mass), _fp(volume),
                                             _ip(material)))

    def add_particles(self, x, v, mass, material=None) -> None:
        """Particles that join the running scene as waiting particles (the
        fluid source, MPM_Example::Add_Fluid_Source); binned from the next
        substep on."""
        x, v, mass = _f32(x), _f32(v), _f32(mass)
        material = _i32(material)
        _check(lib().mpm_add_particles(self._h, 0, mass.shape[0], _fp(x), _fp(v), _fp(mass), _ip(material)))

    def download(self, cap: Optional[int] = None) -> Dict[str, np.ndarray]:
        """Live particles by id; Jp holds an EOS particle's density."""
        if cap is None:
            cap = self.capacity
        out = _particle_arrays(cap, self.dim)
        n = ctypes.c_int64()
        _check(lib().mpm_download_particles_ex(self._h, cap, _fp(out["x"]), _fp(out["v"]), _fp(out["F"]),
                                               _fp(out["Jp"]), _fp(out["mass"]), _fp(out["volume"]),
                                               _ip(out["ids"]), _ip(out["material"]), ctypes.byref(n)))
        return {k: a[:n.value] for k, a in out.items()}

    # -- driver ------------------------------------------------------------
    def initialize(self) -> None:
        _check(lib().mpm_initialize(self._h))

    def substep(self, dt: float, stats: bool = False):
        if stats:
            s = StepStats()
            _check(lib().mpm_substep(self._h, float(dt), ctypes.byref(s)))
            return s.as_dict()
        _check(lib().mpm_substep(self._h, float(dt), None))
        return None

    def advance_to_time(self, time: float, target: float):
        n = ctypes.c_int32()
        t = ctypes.c_double()
        s = StepStats()
        _check(lib().mpm_advance_to_time(self._h, time, target, ctypes.byref(n), ctypes.byref(t), ctypes.byref(s)))
        return t.value, n.value, s.as_dict()

    def max_particle_velocity(self) -> float:
        v = ctypes.c_float()
        _check(lib().mpm_max_particle_velocity(self._h, ctypes.byref(v)))
        return v.value

    def synchronize(self) -> None:
        _check(lib().mpm_synchronize(self._h))

    def run_stage(self, stage: int, dt: float = 0.0) -> None:
        _check(lib().mpm_run_stage(self._h, int(stage), float(dt)))

    def stats(self) -> Dict[str, float]:
        s = StepStats()
        _check(lib().mpm_cg_stats(self._h, ctypes.byref(s)))
        return s.as_dict()

    # -- inspection --------------------------------------------------------
    def _dense_nodes(self) -> int:
        return (self.counts + 2) ** self.dim

    def read_grid(self, channel: int) -> np.ndarray:
        w = 1 if channel == CH_MASS else self.dim
        out = np.zeros(self._dense_nodes() * w, np.float32)
        _check(lib().mpm_read_grid(self._h, channel, _fp(out), out.size))
        shape = (self.counts + 2,) * self.dim + ((w,) if w > 1 else ())
        return out.reshape(shape)

    def read_flags(self) -> np.ndarray:
        out = np.zeros(self._dense_nodes(), np.uint32)
        _check(lib().mpm_read_flags(self._h, out.ctypes.data_as(_UP), out.size))
        return out.reshape((self.counts + 2,) * self.dim)

    def read_block_flags(self):
        """(origins[nblocks, dim], flags[nblocks, 64]) of the active blocks."""
        n = ctypes.c_int32()
        _check(lib().mpm_read_block_flags(self._h, 0, None, None, ctypes.byref(n)))
        origin = np.zeros((max(n.value, 1), self.dim), np.int32)
        flags = np.zeros((max(n.value, 1), 64), np.uint32)
        _check(lib().mpm_read_block_flags(self._h, n.value, origin.ctypes.data_as(_IP),
                                          flags.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), ctypes.byref(n)))
        return origin[:n.value], flags[:n.value]

    def read_bins(self):
        nb = ctypes.c_int32()
        npart = ctypes.c_int64()
        _check(lib().mpm_read_bins(self._h, 0, None, None, 0, None, None, ctypes.byref(nb), ctypes.byref(npart)))
        coords = np.zeros((max(nb.value, 1), self.dim), np.int32)
        counts = np.zeros(max(nb.value, 1), np.int32)
        pblock = np.zeros(max(npart.value, 1), np.int32)
        pids = np.zeros(max(npart.value, 1), np.int32)
        _check(lib().mpm_read_bins(self._h, coords.shape[0], coords.ctypes.data_as(_IP), counts.ctypes.data_as(_IP),
                                   pids.shape[0], pblock.ctypes.data_as(_IP), pids.ctypes.data_as(_IP),
                                   ctypes.byref(nb), ctypes.byref(npart)))
        return coords[:nb.value], counts[:nb.value], pblock[:npart.value], pids[:npart.value]

    # -- timing ------------------------------------------------------------
    def enable_timing(self, on: bool = True) -> None:
        _check(lib().mpm_enable_timing(self._h, 1 if on else 0))

    def time_stage(self, stage: int, reps: int = 20) -> float:
        """Average ms of one launch of a stage kernel (STAGE_P2G, STAGE_CONSTITUTIVE
        or STAGE_G2P) on the last substep's state, back to back between HIP
        events; the particle state is restored afterwards."""
        ms = ctypes.c_float()
        _check(lib().mpm_time_stage(self._h, int(stage), int(reps), ctypes.byref(ms)))
        return ms.value

    def time_cg_operator(self, reps: int = 20) -> float:
        """Average ms of one CG operator launch, timed back to back with HIP events."""
        ms = ctypes.c_float()
        _check(lib().mpm_time_cg_operator(self._h, int(reps), ctypes.byref(ms)))
        return ms.value

    def kernel_times(self) -> Dict[str, Tuple[float, int]]:
        names = kernel_class_names()
        ms = (ctypes.c_double * len(names))()
        n = (ctypes.c_int64 * len(names))()
        _check(lib().mpm_kernel_times(self._h, len(names), ms, n))
        return {names[k]: (ms[k], n[k]) for k in range(len(names))}


class Batch(_LaunchControl):
    """Many independent scenes of one dimension advanced together: every
    kernel of a substep covers all of them, and each runs its own CG."""

    def _nscenes(self) -> int:
        return self.n

    def __init__(self, descs, capacities):
        self.descs = [d.copy() for d in descs]
        self.n = len(self.descs)
        self.dim = self.descs[0].dim
        self.capacities = [int(c) for c in capacities]
        arr = (SceneDesc * self.n)(*self.descs)
        caps = (ctypes.c_int64 * self.n)(*self.capacities)
        h = _H()
        _check(lib().mpm_batch_create(arr, caps, self.n, ctypes.byref(h)))
        self._h = h

    def close(self) -> None:
        if getattr(self, "_h", None):
            lib().mpm_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def upload(self, scene: int, x, v, mass, F=None, Jp=None, volume=None, material=None) -> None:
        x, v, mass = _f32(x), _f32(v), _f32(mass)
        F = None if F is None else _f32(F)
        Jp = None if Jp is None else _f32(Jp)
        volume = None if volume is None else _f32(volume)
        material = _i32(material)
        _check(lib().mpm_batch_upload_ex(self._h, int(scene), mass.shape[0], _fp(x), _fp(v), _fp(F), _fp(Jp),
                                         _fp(mass), _fp(volume), _ip(material)))

    def initialize(self) -> None:
        _check(lib().mpm_initialize(self._h))

    def add_particles(self, scene: int, x, v, mass, material=None) -> None:
        """Waiting particles joining scene `scene` (see Scene.add_particles);
        its download capacity grows by their count."""
        x, v, mass = _f32(x), _f32(v), _f32(mass)
        material = _i32(material)
        _check(lib().mpm_add_particles(self._h, int(scene), mass.shape[0], _fp(x), _fp(v), _fp(mass),
                                       _ip(material)))
        self.capacities[scene] += mass.shape[0]

    def substep(self, dt) -> None:
        d = np.ascontiguousarray(np.broadcast_to(np.asarray(dt, np.float32), (self.n,)), dtype=np.float32)
        _check(lib().mpm_batch_substep(self._h, _fp(d)))

    def max_velocity(self) -> np.ndarray:
        out = np.zeros(self.n, np.float32)
        _check(lib().mpm_batch_max_velocity(self._h, _fp(out)))
        return out

    def stats(self, scene: int) -> Dict[str, float]:
        s = StepStats()
        _check(lib().mpm_batch_stats(self._h, int(scene), ctypes.byref(s)))
        return s.as_dict()

    def download(self, scene: int, cap: Optional[int] = None) -> Dict[str, np.ndarray]:
        d = self.dim
        if cap is None:
            cap = self.capacities[scene]
        out = _particle_arrays(cap, d)
        n = ctypes.c_int64()
        _check(lib().mpm_batch_download_ex(self._h, int(scene), cap, _fp(out["x"]), _fp(out["v"]), _fp(out["F"]),
                                           _fp(out["Jp"]), _fp(out["mass"]), _fp(out["volume"]),
                                           _ip(out["ids"]), _ip(out["material"]), ctypes.byref(n)))
        return {k: a[:n.value] for k, a in out.items()}

    def advance_to_time(self, time, target):
        t = np.ascontiguousarray(np.broadcast_to(np.asarray(time, np.float64), (self.n,)), dtype=np.float64)
        g = np.ascontiguousarray(np.broadcast_to(np.asarray(target, np.float64), (self.n,)), dtype=np.float64)
        t_out = np.zeros(self.n, np.float64)
        steps = np.zeros(self.n, np.int32)
        dp = ctypes.POINTER(ctypes.c_double)
        _check(lib().mpm_batch_advance_to_time(self._h, t.ctypes.data_as(dp), g.ctypes.data_as(dp),
                                               t_out.ctypes.data_as(dp), steps.ctypes.data_as(_IP)))
        return t_out, steps

    def synchronize(self) -> None:
        _check(lib().mpm_synchronize(self._h))

    def enable_timing(self, on: bool = True) -> None:
        _check(lib().mpm_enable_timing(self._h, 1 if on else 0))

    def kernel_times(self) -> Dict[str, Tuple[float, int]]:
        names = kernel_class_names()
        ms = (ctypes.c_double * len(names))()
        n = (ctypes.c_int64 * len(names))()
        _check(lib().mpm_kernel_times(self._h, len(names), ms, n))
        return {names[k]: (ms[k], n[k]) for k in range(len(names))}


class Server:
    """Continuous batching of scene requests on one device arena: submit()
    queues a scene with a frame count, tick() admits what fits and advances
    every resident scene by one frame (mpm_serve_*)."""

    def __init__(self, hbm_budget_bytes: int, max_scenes: int = 64, max_counts: int = 256, dim: int = 3,
                 cg_max_iterations: int = 50, graph: bool = False, keep_results: bool = False,
                 max_tick_substeps: int = 0, tune: bool = False):
        self.dim = dim
        self.keep_results = keep_results
        cfg = ServeConfig(dim, max_scenes, max_counts, cg_max_iterations, int(hbm_budget_bytes),
                          1 if graph else 0, 1 if keep_results else 0, int(max_tick_substeps), 1 if tune else 0)
        h = _H()
        _check(lib().mpm_serve_create(ctypes.byref(cfg), ctypes.byref(h)))
        self._h = h
        self._n = {}

    def close(self) -> None:
        if getattr(self, "_h", None):
            lib().mpm_serve_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def arena(self) -> Dict[str, int]:
        v = [ctypes.c_int64() for _ in range(4)]
        _check(lib().mpm_serve_arena(self._h, *[ctypes.byref(x) for x in v]))
        return dict(zip(("particles", "slots", "blocks", "bytes"), (x.value for x in v)))

    def submit(self, desc: SceneDesc, x, v, mass, frames: int, material=None) -> int:
        x, v, mass = _f32(x), _f32(v), _f32(mass)
        material = _i32(material)
        d = desc.copy()
        rid = ctypes.c_int64()
        _check(lib().mpm_serve_submit_ex(self._h, ctypes.byref(d), mass.shape[0], _fp(x), _fp(v), _fp(mass),
                                         _ip(material), int(frames), ctypes.byref(rid)))
        self._n[rid.value] = mass.shape[0]
        return rid.value

    def tick(self) -> Dict[str, float]:
        st = TickStats()
        _check(lib().mpm_serve_tick(self._h, ctypes.byref(st)))
        return st.as_dict()

    def poll(self, rid: int) -> Dict[str, float]:
        st = RequestStatus()
        _check(lib().mpm_serve_poll(self._h, int(rid), ctypes.byref(st)))
        return st.as_dict()

    def download(self, rid: int) -> Dict[str, np.ndarray]:
        d = self.dim
        cap = self._n[rid]
        out = {"x": np.zeros((cap, d), np.float32), "v": np.zeros((cap, d), np.float32),
               "F": np.zeros((cap, d, d), np.float32), "Jp": np.zeros(cap, np.float32),
               "mass": np.zeros(cap, np.float32), "volume": np.zeros(cap, np.float32),
               "ids": np.zeros(cap, np.int32)}
        n = ctypes.c_int64()
        _check(lib().mpm_serve_download(self._h, int(rid), cap, _fp(out["x"]), _fp(out["v"]), _fp(out["F"]),
                                        _fp(out["Jp"]), _fp(out["mass"]), _fp(out["volume"]),
                                        out["ids"].ctypes.data_as(_IP), ctypes.byref(n)))
        return {k: a[:n.value] for k, a in out.items()}

    def frame_times(self, rid: int) -> np.ndarray:
        """Completion time of each finished frame (server clock, seconds)."""
        n = ctypes.c_int32()
        _check(lib().mpm_serve_frame_times(self._h, int(rid), 0, None, ctypes.byref(n)))
        out = np.zeros(max(n.value, 1), np.float64)
        _check(lib().mpm_serve_frame_times(self._h, int(rid), n.value,
                                           out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), ctypes.byref(n)))
        return out[:n.value]

    def release(self, rid: int) -> None:
        _check(lib().mpm_serve_release(self._h, int(rid)))

    def cancel(self, rid: int) -> None:
        """Withdraws a queued or running request (REQ_CANCELLED)."""
        _check(lib().mpm_serve_cancel(self._h, int(rid)))
        self._n.pop(rid, None)
