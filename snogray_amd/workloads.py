"""Seeded job distributions for batched and served rendering.

A job is an icosphere-soup scene (the 1.31 M-triangle workload's
generator, with its own triangle count and its own Gaussian-lobe
envmap) rendered at its own resolution and sample count:
resolution log-uniform in [64, 1024] per side, spp log-uniform in
[1, 256], triangles log-uniform in [10 k, 2 M], max path length 10.
"""
import math
from dataclasses import dataclass

import numpy as np

from . import abi, scenes


@dataclass(frozen=True)
class SceneSpec:
    """An icosphere soup: n_spheres icospheres of the given level."""
    n_spheres: int
    level: int
    seed: int
    env_w: int = 1024
    env_h: int = 512

    @property
    def triangles(self):
        return self.n_spheres * 20 * 4 ** self.level + 2

    def build(self, aspect=16.0 / 9.0):
        return scenes.icosphere_soup(n_spheres=self.n_spheres, level=self.level, seed=self.seed,
                                     env_size=(self.env_w, self.env_h), aspect=aspect)

    def estimated_bytes(self):
        """Device footprint estimate before the scene is built: octree
        nodes, inline list records and triangle data come to about 200 B
        per triangle on these soups, plus the envmap and its distribution."""
        return int(200 * self.triangles + 24 * self.env_w * self.env_h + (1 << 20))


def scene_for_triangles(tris, seed):
    """Icosphere level and count whose total is close to TRIS."""
    level = int(np.clip(math.floor(math.log(max(tris, 80) / 40.0, 4)), 2, 6))
    n = max(1, int(round(tris / (20 * 4 ** level))))
    return SceneSpec(n_spheres=n, level=level, seed=seed)


@dataclass(frozen=True)
class JobSpec:
    scene: SceneSpec
    width: int
    height: int
    spp: int
    seed: int
    max_path_len: int = 10
    min_path_len: int = 3

    @property
    def samples(self):
        return self.width * self.height * self.spp

    def params(self, **kw):
        return abi.make_params(self.width, self.height, self.spp, max_path_len=self.max_path_len,
                               min_path_len=self.min_path_len, seed=self.seed, **kw)

    def camera(self):
        return scenes.look_at((0.0, 12.0, -34.0), (0.0, 9.0, 0.0), vfov_deg=45.0,
                              aspect=self.width / self.height)


def _log_uniform(rng, lo, hi, size=None):
    return np.exp(rng.uniform(math.log(lo), math.log(hi), size))


def draw_jobs(n, seed, scene_pool=None):
    """N jobs from the distribution.  With SCENE_POOL (a list of
    SceneSpec) each job draws one of those scenes; otherwise each job
    gets its own scene."""
    rng = np.random.default_rng(seed)
    jobs = []
    for i in range(n):
        w = int(round(_log_uniform(rng, 64, 1024)))
        h = int(round(_log_uniform(rng, 64, 1024)))
        spp = int(round(_log_uniform(rng, 1, 256)))
        if scene_pool is None:
            sc = scene_for_triangles(_log_uniform(rng, 1e4, 2e6), seed=int(rng.integers(1 << 30)))
        else:
            sc = scene_pool[int(rng.integers(len(scene_pool)))]
        jobs.append(JobSpec(sc, w, h, spp, seed=int(rng.integers(1 << 31))))
    return jobs


def scene_pool(n, seed):
    rng = np.random.default_rng(seed)
    return [scene_for_triangles(_log_uniform(rng, 1e4, 2e6), seed=int(rng.integers(1 << 30)))
            for _ in range(n)]


def poisson_arrivals(rate, duration, seed):
    """Arrival times (s) of a Poisson process of RATE per second."""
    rng = np.random.default_rng(seed)
    t, out = 0.0, []
    while True:
        t += rng.exponential(1.0 / rate)
        if t >= duration:
            return out
        out.append(t)
