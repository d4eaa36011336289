"""ctypes mirror of include/snoggpu.h.

Plain data structures only; loading the device library lives in
``snogray_amd.native``.  The same descriptors are handed to the CPU
oracle by the tests, so one scene description drives both.
"""
import ctypes as C

import numpy as np

SGR_OK = 0
SGR_MAT_LAMBERT = 1
SGR_MAT_COOK_TORRANCE = 2
SGR_MAT_GLASS = 3
SGR_MAT_MIRROR = 4
SGR_MAT_GLOW = 5
SGR_MAT_PHONG = 6
SGR_MAT_THIN_GLASS = 7
SGR_MAT_STENCIL = 8
SGR_INTEG_PATH, SGR_INTEG_DIRECT, SGR_INTEG_PHOTON = 0, 1, 2
SGR_PHOTON_UNSPEC = 0xFFFFFFFF
SGR_PHOTON_MAX_DEPOSITS = 32
SGR_ACCEL_OCTREE = 0
SGR_ACCEL_TRIV = 1
MISS = 0xFFFFFFFF
SGR_RENDER_KERNEL_TIMING = 1
SGR_RENDER_ONE_CHUNK_IN_FLIGHT = 2
SGR_RENDER_WAVEFRONT_ONLY = 4
SGR_DEFAULT_PATHS_IN_FLIGHT = 1 << 26

f32p = C.POINTER(C.c_float)
u32p = C.POINTER(C.c_uint32)


class Material(C.Structure):
    _fields_ = [("type", C.c_int32), ("color", C.c_float * 3),
                ("gloss_color", C.c_float * 3), ("m", C.c_float),
                ("ior_n", C.c_float), ("ior_k", C.c_float),
                ("underlying", C.c_int32)]


class Mesh(C.Structure):
    _fields_ = [("vertices", f32p), ("num_vertices", C.c_uint32),
                ("triangles", u32p), ("num_triangles", C.c_uint32),
                ("tri_materials", u32p), ("material", C.c_uint32),
                ("axis", C.c_float * 3), ("left_handed", C.c_int32),
                ("normals", f32p)]


class Tripar(C.Structure):
    _fields_ = [("v0", C.c_float * 3), ("e1", C.c_float * 3),
                ("e2", C.c_float * 3), ("parallelogram", C.c_int32),
                ("material", C.c_uint32)]


class Sphere(C.Structure):
    _fields_ = [("center", C.c_float * 3), ("radius", C.c_float),
                ("material", C.c_uint32)]


class Ellipse(C.Structure):
    _fields_ = [("center", C.c_float * 3), ("radius1", C.c_float * 3),
                ("radius2", C.c_float * 3), ("material", C.c_uint32)]


SGR_LOCAL_CYLINDER = 1
SGR_LOCAL_SPHERE2 = 2


class LocalPrim(C.Structure):
    _fields_ = [("kind", C.c_int32), ("local_to_world", C.c_float * 12),
                ("material", C.c_uint32)]


class PointLight(C.Structure):
    _fields_ = [("pos", C.c_float * 3), ("intensity", C.c_float * 3), ("angle", C.c_float),
                ("dir", C.c_float * 3), ("fringe_angle", C.c_float)]


class Envmap(C.Structure):
    _fields_ = [("texels", f32p), ("width", C.c_uint32),
                ("height", C.c_uint32), ("frame", C.c_float * 9)]


class SceneDescC(C.Structure):
    _fields_ = [("meshes", C.POINTER(Mesh)), ("num_meshes", C.c_uint32),
                ("tripars", C.POINTER(Tripar)), ("num_tripars", C.c_uint32),
                ("spheres", C.POINTER(Sphere)), ("num_spheres", C.c_uint32),
                ("materials", C.POINTER(Material)),
                ("num_materials", C.c_uint32),
                ("envmaps", C.POINTER(Envmap)), ("num_envmaps", C.c_uint32),
                ("min_trace", C.c_float), ("accel", C.c_int32),
                ("models", C.c_void_p), ("num_models", C.c_uint32),
                ("instances", C.c_void_p), ("num_instances", C.c_uint32),
                ("ellipses", C.POINTER(Ellipse)), ("num_ellipses", C.c_uint32),
                ("local_prims", C.POINTER(LocalPrim)), ("num_local_prims", C.c_uint32),
                ("point_lights", C.POINTER(PointLight)), ("num_point_lights", C.c_uint32)]


class Model(C.Structure):
    _fields_ = [("meshes", C.POINTER(Mesh)), ("num_meshes", C.c_uint32),
                ("tripars", C.POINTER(Tripar)), ("num_tripars", C.c_uint32),
                ("spheres", C.POINTER(Sphere)), ("num_spheres", C.c_uint32),
                ("ellipses", C.POINTER(Ellipse)), ("num_ellipses", C.c_uint32),
                ("local_prims", C.POINTER(LocalPrim)), ("num_local_prims", C.c_uint32)]


class Instance(C.Structure):
    _fields_ = [("model", C.c_uint32), ("local_to_world", C.c_float * 12)]


class Rays(C.Structure):
    _fields_ = [(k, f32p) for k in ("ox", "oy", "oz", "dx", "dy", "dz",
                                    "t0", "t1")]


class Hits(C.Structure):
    _fields_ = [("prim", u32p), ("t", f32p), ("u", f32p), ("v", f32p), ("inst", u32p)]


class Camera(C.Structure):
    _fields_ = [("pos", C.c_float * 3), ("forward", C.c_float * 3),
                ("up", C.c_float * 3), ("right", C.c_float * 3),
                ("tan_half_fov_x", C.c_float), ("tan_half_fov_y", C.c_float)]


class RenderParams(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32),
                ("spp", C.c_uint32), ("min_path_len", C.c_uint32),
                ("max_path_len", C.c_uint32),
                ("direct_samples", C.c_uint32), ("seed", C.c_uint64),
                ("filter_b", C.c_float), ("filter_c", C.c_float),
                ("filter_radius", C.c_float), ("neg_clamp", C.c_float),
                ("bg_alpha", C.c_float), ("x0", C.c_uint32),
                ("y0", C.c_uint32), ("w", C.c_uint32), ("h", C.c_uint32),
                ("paths_in_flight", C.c_uint32), ("flags", C.c_uint32),
                ("integrator", C.c_uint32), ("light_samples", C.c_uint32),
                ("photon_map", C.c_void_p)]


class PhotonParams(C.Structure):
    _fields_ = [("caustic_photons", C.c_uint32), ("direct_photons", C.c_uint32),
                ("indirect_photons", C.c_uint32), ("use_photons", C.c_uint32),
                ("photon_radius", C.c_float), ("final_gather", C.c_uint32),
                ("fg_samples", C.c_uint32), ("fg_bsdf_samples", C.c_uint32),
                ("fg_photon_samples", C.c_uint32), ("direct_illum", C.c_uint32),
                ("direct_samples", C.c_uint32), ("seed", C.c_uint64)]


class PhotonMapInfo(C.Structure):
    _fields_ = [("photons", C.c_uint64 * 3), ("paths", C.c_uint64 * 3),
                ("scale", C.c_float * 3), ("paths_shot", C.c_uint64),
                ("deposit_overflows", C.c_uint64), ("fg_samples", C.c_uint32),
                ("fg_bsdf_samples", C.c_uint32), ("fg_photon_samples", C.c_uint32),
                ("ms_shoot", C.c_double), ("ms_build", C.c_double)]

    def as_dict(self):
        d = {}
        for k, _ in self._fields_:
            v = getattr(self, k)
            d[k] = list(v) if hasattr(v, "__len__") else v
        return d


def photon_params(caustic=50000, direct=500000, indirect=500000, use_photons=50,
                  radius=0.1, final_gather=True, fg_samples=None, fg_bsdf_samples=None,
                  fg_photon_samples=None, direct_illum=True, direct_samples=16, seed=1):
    """sgr_photon_params with the reference's defaults
    (render/photon-integ.cc:36-148); None leaves a count unspecified."""
    u = lambda x: SGR_PHOTON_UNSPEC if x is None else int(x)
    return PhotonParams(caustic, direct, indirect, use_photons, radius,
                        1 if final_gather else 0, u(fg_samples), u(fg_bsdf_samples),
                        u(fg_photon_samples), 1 if direct_illum else 0, direct_samples, seed)


class Stats(C.Structure):
    _fields_ = [(k, C.c_uint64) for k in (
        "eye_rays", "scene_intersect_calls", "shadow_tests",
        "node_visits_closest", "surf_tests_closest", "node_visits_shadow",
        "surf_tests_shadow", "illum_calls", "samples", "paths_traced")] + \
        [(k, C.c_double) for k in ("ms_render", "ms_extend", "ms_shade",
                                   "ms_shadow", "ms_splat", "ms_other")] + \
        [("extend_launches", C.c_uint32), ("shadow_launches", C.c_uint32),
         ("media_overflows", C.c_uint64), ("node_fetches_closest", C.c_uint64),
         ("node_fetches_shadow", C.c_uint64), ("direct_overflows", C.c_uint64),
         ("partial_overflows", C.c_uint64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class EngineConfig(C.Structure):
    _fields_ = [("path_capacity", C.c_uint64), ("memory_budget", C.c_uint64),
                ("grid_blocks", C.c_uint32), ("use_graph", C.c_uint32),
                ("finish_max", C.c_uint32)]


SGR_JOB_QUEUED, SGR_JOB_ACTIVE, SGR_JOB_DONE = 0, 1, 2


class JobStatus(C.Structure):
    _fields_ = [("state", C.c_int32), ("chunks_done", C.c_uint32),
                ("chunks_total", C.c_uint32), ("steps", C.c_uint32),
                ("t_submit_ms", C.c_double), ("t_start_ms", C.c_double),
                ("t_done_ms", C.c_double), ("footprint", C.c_uint64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class StepInfo(C.Structure):
    _fields_ = [("jobs", C.c_uint32), ("jobs_finished", C.c_uint32),
                ("paths", C.c_uint64), ("ms_device", C.c_double),
                ("max_path_len", C.c_uint32), ("graph", C.c_uint32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class GuideStats(C.Structure):
    _fields_ = [("train_steps", C.c_uint64), ("records", C.c_uint64), ("last_loss", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


GD_NPARAM = 16624
GD_REC = 16


class OctreeStats(C.Structure):
    _fields_ = [("num_nodes", C.c_uint64), ("num_leaf_nodes", C.c_uint64),
                ("num_surfaces", C.c_uint64),
                ("num_dup_surfaces", C.c_uint64), ("max_depth", C.c_uint32),
                ("avg_depth", C.c_float)]


def fptr(a):
    return a.ctypes.data_as(f32p)


def uptr(a):
    return a.ctypes.data_as(u32p)


def make_params(width, height, spp, max_path_len=8, min_path_len=3,
                direct_samples=1, seed=1, window=None, filter_b=1.0 / 3,
                filter_c=1.0 / 3, filter_radius=2.0, neg_clamp=-0.1,
                bg_alpha=1.0, paths_in_flight=0, flags=0, integrator=0, light_samples=0,
                photon_map=None):
    p = RenderParams()
    p.width, p.height, p.spp = width, height, spp
    p.min_path_len, p.max_path_len = min_path_len, max_path_len
    p.direct_samples, p.seed = direct_samples, seed
    p.filter_b, p.filter_c = filter_b, filter_c
    p.filter_radius, p.neg_clamp, p.bg_alpha = filter_radius, neg_clamp, bg_alpha
    if window is not None:
        p.x0, p.y0, p.w, p.h = window
    p.paths_in_flight = paths_in_flight
    p.flags = flags
    p.integrator, p.light_samples = integrator, light_samples
    if photon_map is not None:
        p.integrator = SGR_INTEG_PHOTON
        p.photon_map = photon_map.handle()
    return p


def render_window(p):
    """(x0, y0, w, h, margin) of the pixels a render traces."""
    import math
    margin = int(math.ceil(p.filter_radius - 0.5001))
    if p.w == 0:
        return 0, 0, p.width, p.height, margin
    return p.x0, p.y0, p.w, p.h, margin


class RayBatch:
    """SoA float32 ray batch that keeps its arrays alive."""

    def __init__(self, origins, dirs, t0, t1):
        o = np.ascontiguousarray(origins, dtype=np.float32)
        d = np.ascontiguousarray(dirs, dtype=np.float32)
        self.cols = [np.ascontiguousarray(o[:, 0]), np.ascontiguousarray(o[:, 1]),
                     np.ascontiguousarray(o[:, 2]), np.ascontiguousarray(d[:, 0]),
                     np.ascontiguousarray(d[:, 1]), np.ascontiguousarray(d[:, 2]),
                     np.ascontiguousarray(np.broadcast_to(t0, len(o)), dtype=np.float32),
                     np.ascontiguousarray(np.broadcast_to(t1, len(o)), dtype=np.float32)]
        self.n = len(o)
        self.c = Rays(*[fptr(a) for a in self.cols])


class HitBuffer:
    def __init__(self, n):
        self.prim = np.zeros(n, np.uint32)
        self.t = np.zeros(n, np.float32)
        self.u = np.zeros(n, np.float32)
        self.v = np.zeros(n, np.float32)
        self.inst = np.zeros(n, np.uint32)
        self.c = Hits(uptr(self.prim), fptr(self.t), fptr(self.u), fptr(self.v),
                      uptr(self.inst))
