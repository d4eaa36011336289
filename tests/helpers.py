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


class DeviceMem:
    """Device memory through the HIP runtime's C API (ctypes), for tests
    that hand device pointers to the C ABI."""
    _hip = None

    @classmethod
    def hip(cls):
        if cls._hip is None:
            import ctypes as C
            h = C.CDLL("libamdhip64.so")
            h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
            h.hipFree.argtypes = [C.c_void_p]
            h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
            h.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
            h.hipDeviceSynchronize.argtypes = []
            cls._hip = h
        return cls._hip

    def __init__(self, nbytes):
        import ctypes as C
        self.nbytes = nbytes
        p = C.c_void_p()
        assert self.hip().hipMalloc(C.byref(p), nbytes) == 0, "hipMalloc"
        self.ptr = p.value
        assert self.hip().hipMemset(self.ptr, 0, nbytes) == 0, "hipMemset"

    def to_host(self, shape, dtype=np.float32):
        out = np.empty(shape, dtype)
        assert out.nbytes <= self.nbytes
        assert self.hip().hipDeviceSynchronize() == 0
        assert self.hip().hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, 2) == 0, "hipMemcpy"
        return out

    def free(self):
        if self.ptr:
            self.hip().hipFree(self.ptr)
            self.ptr = None
