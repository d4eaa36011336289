"""Shared test inputs: seeded ray batches and scene builders."""
import numpy as np

from snogray_amd import abi, scenes


def random_rays(lo, hi, n, seed, t0=1e-3, t1=1e3):
    """Origins uniform in the box [lo, hi], directions isotropic."""
    rng = np.random.default_rng(seed)
    o = rng.uniform(lo, hi, (n, 3)).astype(np.float32)
    d = rng.normal(size=(n, 3)).astype(np.float32)
    return abi.RayBatch(o, d, np.float32(t0), np.float32(t1))


def camera_rays(cam, w, h, seed, t1=1e4):
    """Jittered pinhole eye rays (camera/camera.h:322-389), non-unit dirs."""
    rng = np.random.default_rng(seed)
    ys, xs = np.mgrid[0:h, 0:w]
    u = (xs.ravel() + rng.random(w * h)) / w
    v = (h - (ys.ravel() + rng.random(w * h))) / h
    f, r, up = (np.array(cam.forward, np.float64), np.array(cam.right, np.float64),
                np.array(cam.up, np.float64))
    d = (f[None] + r[None] * ((2 * u - 1) * cam.tan_half_fov_x)[:, None]
         + up[None] * ((2 * v - 1) * cam.tan_half_fov_y)[:, None])
    o = np.broadcast_to(np.array(cam.pos, np.float32), d.shape)
    return abi.RayBatch(o, d.astype(np.float32), np.float32(0), np.float32(t1))


def shadow_batch(batch, hits, frac=0.999, miss_t1=50.0):
    """Same rays with t1 just short of the closest hit (or MISS_T1)."""
    t1 = np.where(hits.prim != abi.MISS, hits.t * np.float32(frac),
                  np.float32(miss_t1)).astype(np.float32)
    return abi.RayBatch(np.stack(batch.cols[0:3], 1), np.stack(batch.cols[3:6], 1),
                        batch.cols[6], t1)


def triangle_soup(n, seed, extent=10.0, size=0.5, accel=abi.SGR_ACCEL_OCTREE):
    """N random triangles (one mesh) of side about SIZE in a cube."""
    rng = np.random.default_rng(seed)
    c = rng.uniform(-extent, extent, (n, 1, 3))
    v = (c + rng.normal(scale=size, size=(n, 3, 3))).reshape(-1, 3)
    sd = scenes.SceneDesc(accel=accel)
    m = sd.lambert((0.5, 0.5, 0.5))
    sd.mesh(v, np.arange(3 * n, dtype=np.uint32).reshape(-1, 3), m)
    return sd


def small_soup(level=4, n_spheres=2, seed=7):
    """Reduced icosphere soup (same generator as the full workload)."""
    return scenes.icosphere_soup(n_spheres=n_spheres, level=level, seed=seed,
                                 env_size=(128, 64))
