"""Seeded scene mixes of the batched and served forms.

Batched form: 256 independent 3D snow scenes, each a random blob (sphere)
or slab (box), with grid size drawn from 32^3 to 128^3 and particle count
log-uniform between 2k and 200k; scene s uses seed 1000 + s.

Served form: a Poisson arrival stream of snow-impact ("two_balls") and
"avalanche" requests, particle count log-uniform between 10k and 2M, grid
up to 256^3, 60 to 240 frames each.

Both only describe scenes; the particles come from the library's sized
scene presets (mpm_scene_generate_sized), so the same (name, counts, seed,
n) always gives the same particles.
"""
from dataclasses import dataclass
from typing import List

import numpy as np


@dataclass
class SceneSpec:
    name: str     # preset name
    counts: int   # cells per axis
    seed: int
    n: int        # particles

    def generate(self):
        from . import scene_generate
        return scene_generate(self.name, self.counts, self.seed, self.n)


def batch_scenes(count: int = 256, first_seed: int = 1000, n_min: int = 2000, n_max: int = 200_000,
                 counts_min: int = 32, counts_max: int = 128) -> List[SceneSpec]:
    """Scene s: seed first_seed + s; blob or slab with equal odds; counts a
    multiple of 4 uniform in [counts_min, counts_max]; particle count
    log-uniform in [n_min, n_max]."""
    out = []
    for s in range(count):
        seed = first_seed + s
        r = np.random.default_rng(seed)
        name = "blob" if r.random() < 0.5 else "slab"
        counts = 4 * int(r.integers(counts_min // 4, counts_max // 4 + 1))
        n = int(round(np.exp(r.uniform(np.log(n_min), np.log(n_max)))))
        out.append(SceneSpec(name, counts, seed, n))
    return out


@dataclass
class Request:
    rid: int
    arrival: float   # seconds after the stream starts
    spec: SceneSpec
    frames: int


def serve_grid(n: int, max_counts: int = 256) -> int:
    """Grid of a served request: the smallest of 64^3, 128^3, 256^3 whose
    template keeps about 8 particles per cell or fewer."""
    for c in (64, 128, 256):
        if c >= max_counts or n <= 8 * 0.0075 * c ** 3:
            return min(c, max_counts)
    return max_counts


def serve_requests(count: int, rate_hz: float, seed: int = 7, n_min: int = 10_000, n_max: int = 2_000_000,
                   frames_min: int = 60, frames_max: int = 240, max_counts: int = 256) -> List[Request]:
    """Poisson arrivals at rate_hz; each request a snow-impact or avalanche
    template with log-uniform particle count and uniform frame count."""
    r = np.random.default_rng(seed)
    t = 0.0
    out = []
    for k in range(count):
        t += float(r.exponential(1.0 / rate_hz))
        name = "two_balls" if r.random() < 0.5 else "avalanche"
        n = int(round(np.exp(r.uniform(np.log(n_min), np.log(n_max)))))
        frames = int(r.integers(frames_min, frames_max + 1))
        out.append(Request(k, t, SceneSpec(name, serve_grid(n, max_counts), 5000 + k, n), frames))
    return out
