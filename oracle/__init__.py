"""ctypes binding of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's CPU-baseline leg, never by the product.  See mpm_oracle.h.
"""
from __future__ import annotations

import ctypes
import os
from typing import Dict, Optional

import numpy as np

from non_ficks_diffusion_amd import SceneDesc, Material

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")

DOT_FP64, SVD_FP64, JP_SCALAR = 1, 2, 4

_FP = ctypes.POINTER(ctypes.c_float)
_DP = ctypes.POINTER(ctypes.c_double)
_H = ctypes.c_void_p

_SIG = {
    "oracle_create": (_H, [ctypes.POINTER(SceneDesc), ctypes.c_int32, ctypes.c_int32, ctypes.c_int32]),
    "oracle_destroy": (None, [_H]),
    "oracle_set_particles": (None, [_H, ctypes.c_int64, _FP, _FP, _FP, _FP, _FP, _FP]),
    "oracle_set_particles_ex": (None, [_H, ctypes.c_int64, _FP, _FP, _FP, _FP, _FP, _FP, ctypes.POINTER(ctypes.c_int32)]),
    "oracle_add_particles": (None, [_H, ctypes.c_int64, _FP, _FP, _FP, ctypes.POINTER(ctypes.c_int32)]),
    "oracle_num_particles": (ctypes.c_int64, [_H]),
    "oracle_num_simulated": (ctypes.c_int64, [_H]),
    "oracle_get_particles": (None, [_H, _DP, _DP, _DP, _DP, _DP, _DP, ctypes.POINTER(ctypes.c_uint8)]),
    "oracle_initialize": (None, [_H]),
    "oracle_advance_step": (None, [_H, ctypes.c_double]),
    "oracle_run_stage": (None, [_H, ctypes.c_int32, ctypes.c_double]),
    "oracle_populate_simulated": (None, [_H]),
    "oracle_max_particle_velocity": (ctypes.c_double, [_H]),
    "oracle_advance_to_target_time": (ctypes.c_double, [_H, ctypes.c_double, ctypes.c_double,
                                                        ctypes.POINTER(ctypes.c_int32)]),
    "oracle_get_grid": (None, [_H, ctypes.c_int32, _DP]),
    "oracle_get_flags": (None, [_H, ctypes.POINTER(ctypes.c_uint32)]),
    "oracle_apply_operator": (None, [_H, ctypes.c_double, _DP, _DP]),
    "oracle_inner_product": (ctypes.c_double, [_H, _DP, _DP]),
    "oracle_set_grid": (None, [_H, ctypes.c_int32, _DP]),
    "oracle_cg_stats": (None, [_H, ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32),
                               ctypes.POINTER(ctypes.c_int32), _DP]),
    "oracle_bbox_cells": (None, [_H, ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32)]),
    "oracle_weights": (None, [ctypes.c_int32, ctypes.c_int32, _DP, ctypes.POINTER(ctypes.c_int32), _DP, _DP]),
    "oracle_svd": (None, [ctypes.c_int32, _DP, _DP, _DP, _DP]),
    "oracle_constitutive": (None, [ctypes.c_int32, ctypes.POINTER(Material), _DP, ctypes.c_double, _DP, _DP,
                                   _DP, _DP, _DP]),
    "oracle_dP": (None, [ctypes.c_int32, _DP, _DP, ctypes.c_double, ctypes.c_double, _DP, _DP]),
}

_lib: Optional[ctypes.CDLL] = None


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: run make -C oracle")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIG.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _dp(a):
    return None if a is None else a.ctypes.data_as(_DP)


def _fp(a):
    return None if a is None else a.ctypes.data_as(_FP)


def _f32(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float32)


class Oracle:
    """Reference-faithful CPU MPM (see mpm_oracle.h)."""

    def __init__(self, desc: SceneDesc, precision: int = 4, threads: int = 1, options: int = 0):
        self.desc = desc.copy()
        self.dim = desc.dim
        self.counts = desc.counts
        self._h = lib().oracle_create(ctypes.byref(self.desc), precision, threads, options)
        if not self._h:
            raise ValueError("oracle_create rejected the scene")

    def close(self):
        if getattr(self, "_h", None):
            lib().oracle_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_particles(self, x, v, mass, F=None, Jp=None, volume=None, material=None):
        """material: entry of the scene's material table per particle (None: all 0);
        an EOS particle's Jp is its density (None: its material's)."""
        x, v, mass, F, Jp, volume = map(_f32, (x, v, mass, F, Jp, volume))
        mat = None if material is None else np.ascontiguousarray(material, dtype=np.int32)
        mp = None if mat is None else mat.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
        lib().oracle_set_particles_ex(self._h, mass.shape[0], _fp(x), _fp(v), _fp(F), _fp(Jp), _fp(mass), _fp(volume), mp)
        self._keep_mat = mat

    def add_particles(self, x, v, mass, material=None):
        """Waiting particles that join between substeps (Add_Fluid_Source)."""
        x, v, mass = map(_f32, (x, v, mass))
        mat = None if material is None else np.ascontiguousarray(material, dtype=np.int32)
        mp = None if mat is None else mat.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
        lib().oracle_add_particles(self._h, mass.shape[0], _fp(x), _fp(v), _fp(mass), mp)

    def particles(self, only_valid: bool = True) -> Dict[str, np.ndarray]:
        n = lib().oracle_num_particles(self._h)
        d = self.dim
        out = {"x": np.zeros((n, d)), "v": np.zeros((n, d)), "F": np.zeros((n, d, d)), "Jp": np.zeros(n),
               "mass": np.zeros(n), "volume": np.zeros(n)}
        valid = np.zeros(n, np.uint8)
        lib().oracle_get_particles(self._h, _dp(out["x"]), _dp(out["v"]), _dp(out["F"]), _dp(out["Jp"]),
                                   _dp(out["mass"]), _dp(out["volume"]), valid.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)))
        out["ids"] = np.arange(n, dtype=np.int32)
        if only_valid:
            keep = valid.astype(bool)
            out = {k: a[keep] for k, a in out.items()}
        return out

    def initialize(self):
        lib().oracle_initialize(self._h)

    def advance_step(self, dt: float):
        lib().oracle_advance_step(self._h, dt)

    def run_stage(self, stage: int, dt: float = 0.0):
        lib().oracle_run_stage(self._h, stage, dt)

    def max_particle_velocity(self) -> float:
        return lib().oracle_max_particle_velocity(self._h)

    def advance_to_target_time(self, time: float, target: float):
        n = ctypes.c_int32()
        t = lib().oracle_advance_to_target_time(self._h, time, target, ctypes.byref(n))
        return t, n.value

    def grid(self, channel: int) -> np.ndarray:
        w = 1 if channel == 0 else self.dim
        out = np.zeros((self.counts + 2) ** self.dim * w)
        lib().oracle_get_grid(self._h, channel, _dp(out))
        return out.reshape((self.counts + 2,) * self.dim + ((w,) if w > 1 else ()))

    def flags(self) -> np.ndarray:
        out = np.zeros((self.counts + 2) ** self.dim, np.uint32)
        lib().oracle_get_flags(self._h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)))
        return out.reshape((self.counts + 2,) * self.dim)

    def set_grid(self, channel: int, values: np.ndarray) -> None:
        v = np.ascontiguousarray(values, dtype=np.float64)
        lib().oracle_set_grid(self._h, channel, _dp(v))

    def apply_operator(self, dt: float, x: np.ndarray) -> np.ndarray:
        x = np.ascontiguousarray(x, dtype=np.float64)
        q = np.zeros_like(x)
        lib().oracle_apply_operator(self._h, dt, _dp(x), _dp(q))
        return q

    def inner_product(self, a: np.ndarray, b: np.ndarray) -> float:
        a = np.ascontiguousarray(a, dtype=np.float64)
        b = np.ascontiguousarray(b, dtype=np.float64)
        return lib().oracle_inner_product(self._h, _dp(a), _dp(b))

    def cg_stats(self):
        it, m, c = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
        r = ctypes.c_double()
        lib().oracle_cg_stats(self._h, ctypes.byref(it), ctypes.byref(m), ctypes.byref(c), ctypes.byref(r))
        return {"cg_iterations": it.value, "cg_multiplies": m.value, "cg_converged": c.value, "cg_residual": r.value}


def weights(dim: int, counts: int, X):
    X = np.ascontiguousarray(X, dtype=np.float64)
    cell = (ctypes.c_int32 * dim)()
    w = np.zeros((3, dim))
    dw = np.zeros((3, dim))
    lib().oracle_weights(dim, counts, _dp(X), cell, _dp(w), _dp(dw))
    return np.array(list(cell)), w, dw


def svd(F):
    F = np.ascontiguousarray(F, dtype=np.float64)
    d = F.shape[0]
    U = np.zeros((d, d))
    V = np.zeros((d, d))
    S = np.zeros(d)
    lib().oracle_svd(d, _dp(F), _dp(U), _dp(S), _dp(V))
    return U, S, V


def constitutive(material: Material, F, Jp: float):
    F = np.ascontiguousarray(F, dtype=np.float64)
    d = F.shape[0]
    Fo, R, P = np.zeros((d, d)), np.zeros((d, d)), np.zeros((d, d))
    Jo, psi = ctypes.c_double(), ctypes.c_double()
    lib().oracle_constitutive(d, ctypes.byref(material), _dp(F), Jp, _dp(Fo), ctypes.byref(Jo), _dp(R), _dp(P),
                              ctypes.byref(psi))
    return Fo, Jo.value, R, P, psi.value


def dP(F, R, lam: float, mu: float, dF):
    F, R, dF = (np.ascontiguousarray(a, dtype=np.float64) for a in (F, R, dF))
    out = np.zeros_like(F)
    lib().oracle_dP(F.shape[0], _dp(F), _dp(R), lam, mu, _dp(dF), _dp(out))
    return out
