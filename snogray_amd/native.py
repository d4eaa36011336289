"""ctypes front end of libsnoggpu.so (include/snoggpu.h).

There is no fallback: if the library is missing or fails to load, every
entry point raises.  Build it with ``python -m snogray_amd.build``.
"""
import atexit
import ctypes as C
import os
import weakref

import numpy as np

from . import abi

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsnoggpu.so")

EXPORTS = [
    "sgr_last_error", "sgr_ctx_create", "sgr_ctx_destroy", "sgr_ctx_memory",
    "sgr_scene_create", "sgr_scene_destroy", "sgr_scene_octree_stats",
    "sgr_octree_export", "sgr_octree_build", "sgr_scene_bounds", "sgr_scene_footprint",
    "sgr_trace_closest", "sgr_trace_occluded", "sgr_render",
    "sgr_render_samples", "sgr_finalize_image", "sgr_render_noalloc",
    "sgr_image_write", "sgr_image_read", "sgr_image_recover",
    "sgr_mesh_file_open", "sgr_mesh_file_copy", "sgr_mesh_file_parts", "sgr_mesh_file_close",
    "sgr_engine_configure", "sgr_job_footprint", "sgr_job_submit", "sgr_engine_step",
    "sgr_job_poll", "sgr_job_result", "sgr_job_release", "sgr_engine_stats",
    "sgr_render_batch",
    "sgr_guide_create", "sgr_guide_destroy", "sgr_guide_get_params", "sgr_guide_set_params",
    "sgr_guide_forward", "sgr_guide_train_batch", "sgr_guide_stats_get", "sgr_render_guided",
]

_lib = None


class SgrError(RuntimeError):
    pass


def lib():
    """Load libsnoggpu.so (raises if it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise SgrError(f"{LIB_PATH} is not built; run python -m snogray_amd.build")
    L = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    L.sgr_last_error.restype = C.c_char_p
    L.sgr_ctx_create.argtypes = [C.c_int, C.POINTER(vp)]
    L.sgr_ctx_destroy.argtypes = [vp]
    L.sgr_ctx_memory.argtypes = [vp, C.POINTER(C.c_uint64)]
    L.sgr_scene_create.argtypes = [vp, C.POINTER(abi.SceneDescC), C.POINTER(vp)]
    L.sgr_scene_destroy.argtypes = [vp]
    L.sgr_scene_octree_stats.argtypes = [vp, C.POINTER(abi.OctreeStats)]
    L.sgr_octree_export.argtypes = [vp, abi.u32p, C.POINTER(C.c_uint64), abi.u32p,
                                    C.POINTER(C.c_uint64), abi.f32p]
    L.sgr_octree_build.argtypes = [C.POINTER(abi.SceneDescC), abi.u32p,
                                   C.POINTER(C.c_uint64), abi.u32p,
                                   C.POINTER(C.c_uint64), abi.f32p,
                                   C.POINTER(abi.OctreeStats)]
    L.sgr_scene_bounds.argtypes = [vp, abi.f32p, abi.f32p, abi.f32p]
    L.sgr_scene_footprint.argtypes = [vp, C.POINTER(C.c_uint64)]
    L.sgr_trace_closest.argtypes = [vp, C.POINTER(abi.Rays), C.c_size_t,
                                    C.POINTER(abi.Hits), C.POINTER(abi.Stats)]
    L.sgr_trace_occluded.argtypes = [vp, C.POINTER(abi.Rays), C.c_size_t,
                                     C.POINTER(C.c_uint8), C.POINTER(abi.Stats)]
    L.sgr_render.argtypes = [vp, C.POINTER(abi.Camera), C.POINTER(abi.RenderParams),
                             abi.f32p, C.c_int, C.POINTER(abi.Stats)]
    L.sgr_render_samples.argtypes = [vp, C.POINTER(abi.Camera),
                                     C.POINTER(abi.RenderParams), abi.f32p,
                                     C.POINTER(abi.Stats)]
    L.sgr_image_write.argtypes = [C.c_char_p, C.c_char_p, abi.f32p, C.c_uint32, C.c_uint32]
    L.sgr_image_read.argtypes = [C.c_char_p, C.c_char_p, abi.f32p, C.POINTER(C.c_uint32),
                                 C.POINTER(C.c_uint32)]
    L.sgr_image_recover.argtypes = [C.c_char_p, C.c_char_p, abi.f32p, C.c_uint32, C.c_uint32,
                                    C.POINTER(C.c_uint32)]
    L.sgr_mesh_file_open.argtypes = [C.c_char_p, C.c_char_p, C.POINTER(vp),
                                     C.POINTER(C.c_uint64), C.POINTER(C.c_uint64),
                                     C.POINTER(C.c_int)]
    L.sgr_mesh_file_copy.argtypes = [vp, abi.f32p, abi.u32p, abi.f32p]
    L.sgr_mesh_file_parts.argtypes = [vp, C.POINTER(C.c_int), C.POINTER(C.c_uint32),
                                      abi.u32p, abi.f32p]
    L.sgr_cylinder_xform.argtypes = [abi.f32p, abi.f32p, C.c_float, abi.f32p]
    L.sgr_photon_params_default.argtypes = [C.POINTER(abi.PhotonParams)]
    L.sgr_photon_params_default.restype = None
    L.sgr_photon_map_create.argtypes = [vp, C.POINTER(abi.PhotonParams), C.POINTER(vp)]
    L.sgr_photon_map_info_get.argtypes = [vp, C.POINTER(abi.PhotonMapInfo)]
    L.sgr_photon_map_export.argtypes = [vp, C.c_int, abi.f32p, abi.f32p, abi.f32p]
    L.sgr_photon_map_destroy.argtypes = [vp]
    L.sgr_mesh_file_close.argtypes = [vp]
    L.sgr_render_noalloc.argtypes = [vp, C.POINTER(abi.Camera), C.POINTER(abi.RenderParams),
                                     abi.f32p, C.POINTER(abi.Stats)]
    L.sgr_finalize_image.argtypes = [abi.f32p, C.c_uint32, C.c_uint32, C.c_float,
                                     C.c_float, abi.f32p]
    L.sgr_engine_configure.argtypes = [vp, C.POINTER(abi.EngineConfig)]
    L.sgr_job_footprint.argtypes = [C.POINTER(abi.RenderParams), C.c_int,
                                    C.POINTER(C.c_uint64)]
    L.sgr_job_submit.argtypes = [vp, vp, C.POINTER(abi.Camera), C.POINTER(abi.RenderParams),
                                 abi.f32p, C.POINTER(C.c_uint64)]
    L.sgr_engine_step.argtypes = [vp, C.POINTER(abi.StepInfo)]
    L.sgr_job_poll.argtypes = [vp, C.c_uint64, C.POINTER(abi.JobStatus)]
    L.sgr_job_result.argtypes = [vp, C.c_uint64, abi.f32p, C.c_int]
    L.sgr_job_release.argtypes = [vp, C.c_uint64]
    L.sgr_engine_stats.argtypes = [vp, C.POINTER(abi.Stats)]
    L.sgr_render_batch.argtypes = [vp, C.c_uint32, C.POINTER(vp), C.POINTER(abi.Camera),
                                   C.POINTER(abi.RenderParams), C.POINTER(abi.f32p),
                                   C.c_int, C.POINTER(abi.Stats)]
    L.sgr_guide_create.argtypes = [vp, C.c_uint64, C.POINTER(vp)]
    L.sgr_guide_destroy.argtypes = [vp]
    L.sgr_guide_get_params.argtypes = [vp, abi.f32p, C.POINTER(C.c_uint64)]
    L.sgr_guide_set_params.argtypes = [vp, abi.f32p, C.c_uint64]
    L.sgr_guide_forward.argtypes = [vp, abi.f32p, C.c_uint64, abi.f32p, C.c_int]
    L.sgr_guide_train_batch.argtypes = [vp, abi.f32p, C.c_uint64, C.c_float, abi.f32p, abi.f32p]
    L.sgr_guide_stats_get.argtypes = [vp, C.POINTER(abi.GuideStats)]
    L.sgr_render_guided.argtypes = [vp, C.POINTER(abi.Camera), C.POINTER(abi.RenderParams), vp,
                                    C.c_int, abi.f32p, C.c_int, C.POINTER(abi.Stats)]
    _lib = L
    return L


def check(rc, what):
    if rc != 0:
        msg = lib().sgr_last_error()
        raise SgrError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")


_live_contexts = weakref.WeakSet()


@atexit.register
def _close_all():
    """Release device objects while the HIP runtime is still up."""
    for ctx in list(_live_contexts):
        ctx.close()


class Context:
    """One device context (HIP stream + work buffers) of libsnoggpu.
    Closing it first closes every scene created on it."""

    def __init__(self, device=0):
        h = C.c_void_p()
        check(lib().sgr_ctx_create(device, C.byref(h)), "sgr_ctx_create")
        self.h = h
        self.scenes = weakref.WeakSet()
        self.engines = weakref.WeakSet()
        self.guides = weakref.WeakSet()
        _live_contexts.add(self)

    def close(self):
        if getattr(self, "h", None):
            for eng in list(self.engines):
                eng.release_all()
            for g in list(self.guides):
                g.close()
            for sc in list(self.scenes):
                sc.close()
            check(lib().sgr_ctx_destroy(self.h), "sgr_ctx_destroy")
            self.h = None

    def __del__(self):
        self.close()

    def memory(self):
        b = C.c_uint64()
        check(lib().sgr_ctx_memory(self.h, C.byref(b)), "sgr_ctx_memory")
        return b.value


class Scene:
    """Device-resident scene built from a SceneDesc."""

    def __init__(self, ctx, desc):
        self.ctx = ctx
        self.desc = desc
        h = C.c_void_p()
        check(lib().sgr_scene_create(ctx.h, C.byref(desc.c()), C.byref(h)),
              "sgr_scene_create")
        self.h = h
        ctx.scenes.add(self)

    def close(self):
        if getattr(self, "h", None):
            for m in list(getattr(self, "photon_maps", ())):
                m.close()
            lib().sgr_scene_destroy(self.h)
            self.h = None

    def photon_map(self, params=None):
        """Shoot this scene's photons and build its maps
        (sgr_photon_map_create); PARAMS: abi.photon_params(...)."""
        return PhotonMap(self, params if params is not None else abi.photon_params())

    def __del__(self):
        self.close()

    def octree_stats(self):
        st = abi.OctreeStats()
        check(lib().sgr_scene_octree_stats(self.h, C.byref(st)), "octree_stats")
        return {k: getattr(st, k) for k, _ in st._fields_}

    def octree(self):
        nn, ns = C.c_uint64(), C.c_uint64()
        check(lib().sgr_octree_export(self.h, None, C.byref(nn), None, C.byref(ns), None),
              "octree_export")
        nodes = np.zeros((nn.value, 4), np.uint32)
        surfs = np.zeros(ns.value, np.uint32)
        osz = np.zeros(4, np.float32)
        check(lib().sgr_octree_export(self.h, abi.uptr(nodes), C.byref(nn),
                                      abi.uptr(surfs), C.byref(ns), abi.fptr(osz)),
              "octree_export")
        return nodes, surfs, osz

    def bounds(self):
        lo, hi, hz = np.zeros(3, np.float32), np.zeros(3, np.float32), np.zeros(1, np.float32)
        check(lib().sgr_scene_bounds(self.h, abi.fptr(lo), abi.fptr(hi), abi.fptr(hz)),
              "scene_bounds")
        return lo, hi, float(hz[0])

    def footprint(self):
        b = C.c_uint64()
        check(lib().sgr_scene_footprint(self.h, C.byref(b)), "scene_footprint")
        return b.value

    def trace_closest(self, batch):
        hb = abi.HitBuffer(batch.n)
        st = abi.Stats()
        check(lib().sgr_trace_closest(self.h, C.byref(batch.c), batch.n, C.byref(hb.c),
                                      C.byref(st)), "sgr_trace_closest")
        return hb, st.as_dict()

    def trace_occluded(self, batch):
        occ = np.zeros(batch.n, np.uint8)
        st = abi.Stats()
        check(lib().sgr_trace_occluded(self.h, C.byref(batch.c), batch.n,
                                       occ.ctypes.data_as(C.POINTER(C.c_uint8)),
                                       C.byref(st)), "sgr_trace_occluded")
        return occ, st.as_dict()

    def render(self, cam, params, accum=None):
        """Render into host accumulators (H, W, 5) float32; returns (accum, stats)."""
        if accum is None:
            accum = np.zeros((params.height, params.width, 5), np.float32)
        assert accum.dtype == np.float32 and accum.flags.c_contiguous
        st = abi.Stats()
        check(lib().sgr_render(self.h, C.byref(cam), C.byref(params), abi.fptr(accum), 0,
                               C.byref(st)), "sgr_render")
        return accum, st.as_dict()

    def render_device(self, cam, params, accum_ptr):
        """Render into device accumulators at accum_ptr (int address)."""
        st = abi.Stats()
        check(lib().sgr_render(self.h, C.byref(cam), C.byref(params),
                               C.cast(C.c_void_p(accum_ptr), abi.f32p), 1, C.byref(st)),
              "sgr_render")
        return st.as_dict()

    def render_guided(self, cam, params, guide, train=True, accum=None):
        """sgr_render_guided into host accumulators; returns (accum, stats)."""
        if accum is None:
            accum = np.zeros((params.height, params.width, 5), np.float32)
        st = abi.Stats()
        check(lib().sgr_render_guided(self.h, C.byref(cam), C.byref(params), guide.h,
                                      int(train), abi.fptr(accum), 0, C.byref(st)),
              "sgr_render_guided")
        return accum, st.as_dict()

    def render_guided_device(self, cam, params, guide, accum_ptr, train=True):
        st = abi.Stats()
        check(lib().sgr_render_guided(self.h, C.byref(cam), C.byref(params), guide.h,
                                      int(train), C.cast(C.c_void_p(accum_ptr), abi.f32p), 1,
                                      C.byref(st)), "sgr_render_guided")
        return st.as_dict()

    def render_samples(self, cam, params):
        x0, y0, w, h, m = abi.render_window(params)
        out = np.zeros(((h + 2 * m) * (w + 2 * m) * params.spp, 6), np.float32)
        st = abi.Stats()
        check(lib().sgr_render_samples(self.h, C.byref(cam), C.byref(params),
                                       abi.fptr(out), C.byref(st)), "sgr_render_samples")
        return out, st.as_dict()


class PhotonMap:
    """Photon maps of a scene (sgr_photon_map_*); pass it to
    abi.make_params(photon_map=...) to render with the photon integrator."""

    def __init__(self, scene, params):
        self.scene = scene
        self.params = params
        h = C.c_void_p()
        check(lib().sgr_photon_map_create(scene.h, C.byref(params), C.byref(h)),
              "sgr_photon_map_create")
        self.h = h
        if not hasattr(scene, "photon_maps"):
            scene.photon_maps = weakref.WeakSet()
        scene.photon_maps.add(self)

    def handle(self):
        return self.h.value

    def info(self):
        i = abi.PhotonMapInfo()
        check(lib().sgr_photon_map_info_get(self.h, C.byref(i)), "sgr_photon_map_info_get")
        return i.as_dict()

    def export(self, which):
        """Set WHICH (0 direct, 1 caustic, 2 indirect) in kd-tree order:
        (pos_axis (n, 4) float32 with the split axis's bits in column 3,
        power (n, 3), dir (n, 3))."""
        n = self.info()["photons"][which]
        pa = np.zeros((n, 4), np.float32)
        pw = np.zeros((n, 3), np.float32)
        dr = np.zeros((n, 3), np.float32)
        check(lib().sgr_photon_map_export(self.h, which, abi.fptr(pa), abi.fptr(pw), abi.fptr(dr)),
              "sgr_photon_map_export")
        return pa, pw, dr

    def close(self):
        if getattr(self, "h", None):
            lib().sgr_photon_map_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


class Engine:
    """The context's multi-job render engine (sgr_job_* calls).

    Jobs are sgr_render calls queued on one context; ``step`` admits what
    fits the memory budget and runs the next sample chunk of every active
    job that fits the path capacity in one launch sequence."""

    def __init__(self, ctx, path_capacity=0, memory_budget=0, grid_blocks=0, use_graph=True,
                 finish_max=0):
        self.ctx = ctx
        self.jobs = set()
        ctx.engines.add(self)
        self.configure(path_capacity, memory_budget, grid_blocks, use_graph, finish_max)

    def configure(self, path_capacity=0, memory_budget=0, grid_blocks=0, use_graph=True,
                  finish_max=0):
        cfg = abi.EngineConfig(path_capacity, memory_budget, grid_blocks, int(bool(use_graph)),
                               finish_max)
        check(lib().sgr_engine_configure(self.ctx.h, C.byref(cfg)), "sgr_engine_configure")
        self.cfg = cfg

    @staticmethod
    def footprint(params, caller_accum=False):
        b = C.c_uint64()
        check(lib().sgr_job_footprint(C.byref(params), int(caller_accum), C.byref(b)),
              "sgr_job_footprint")
        return b.value

    def submit(self, scene, cam, params, accum_ptr=None):
        jid = C.c_uint64()
        acc = C.cast(C.c_void_p(accum_ptr), abi.f32p) if accum_ptr else None
        check(lib().sgr_job_submit(self.ctx.h, scene.h, C.byref(cam), C.byref(params), acc,
                                   C.byref(jid)), "sgr_job_submit")
        self.jobs.add(jid.value)
        return jid.value

    def step(self):
        si = abi.StepInfo()
        check(lib().sgr_engine_step(self.ctx.h, C.byref(si)), "sgr_engine_step")
        return si.as_dict()

    def poll(self, jid):
        s = abi.JobStatus()
        check(lib().sgr_job_poll(self.ctx.h, jid, C.byref(s)), "sgr_job_poll")
        return s.as_dict()

    def result(self, jid, width, height):
        out = np.zeros((height, width, 5), np.float32)
        check(lib().sgr_job_result(self.ctx.h, jid, abi.fptr(out), 0), "sgr_job_result")
        return out

    def result_device(self, jid, dst_ptr):
        check(lib().sgr_job_result(self.ctx.h, jid, C.cast(C.c_void_p(dst_ptr), abi.f32p), 1),
              "sgr_job_result")

    def release(self, jid):
        check(lib().sgr_job_release(self.ctx.h, jid), "sgr_job_release")
        self.jobs.discard(jid)

    def release_all(self):
        for jid in list(self.jobs):
            self.release(jid)

    def stats(self):
        st = abi.Stats()
        check(lib().sgr_engine_stats(self.ctx.h, C.byref(st)), "sgr_engine_stats")
        return st.as_dict()

    def run(self):
        """Step until no job has work left; returns the step infos."""
        steps = []
        while True:
            si = self.step()
            if si["jobs"] == 0:
                return steps
            steps.append(si)


class Guide:
    """The learned directional proposal (sgr_guide_*)."""

    def __init__(self, ctx, seed=1):
        self.ctx = ctx
        h = C.c_void_p()
        check(lib().sgr_guide_create(ctx.h, seed, C.byref(h)), "sgr_guide_create")
        self.h = h
        ctx.guides.add(self)

    def close(self):
        if getattr(self, "h", None):
            lib().sgr_guide_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def params(self):
        out = np.zeros(abi.GD_NPARAM, np.float32)
        n = C.c_uint64()
        check(lib().sgr_guide_get_params(self.h, abi.fptr(out), C.byref(n)), "sgr_guide_get_params")
        return out

    def set_params(self, w):
        w = np.ascontiguousarray(w, np.float32)
        check(lib().sgr_guide_set_params(self.h, abi.fptr(w), len(w)), "sgr_guide_set_params")

    def forward(self, in6, lobes=False):
        x = np.ascontiguousarray(in6, np.float32).reshape(-1, 6)
        out = np.zeros((len(x), 40), np.float32)
        check(lib().sgr_guide_forward(self.h, abi.fptr(x), len(x), abi.fptr(out), int(lobes)),
              "sgr_guide_forward")
        return out

    def train_batch(self, rec, lr=0.0):
        """rec: (n, 16) records.  Returns (gradient, (loss sum, weight sum))."""
        r = np.ascontiguousarray(rec, np.float32).reshape(-1, abi.GD_REC)
        g = np.zeros(abi.GD_NPARAM, np.float32)
        loss = np.zeros(2, np.float32)
        check(lib().sgr_guide_train_batch(self.h, abi.fptr(r), len(r), lr, abi.fptr(g),
                                          abi.fptr(loss)), "sgr_guide_train_batch")
        return g, (float(loss[0]), float(loss[1]))

    def stats(self):
        st = abi.GuideStats()
        check(lib().sgr_guide_stats_get(self.h, C.byref(st)), "sgr_guide_stats_get")
        return st.as_dict()


def render_batch(ctx, jobs):
    """sgr_render_batch over [(scene, cam, params)]: host accumulators
    (H, W, 5) per job and the batch's counters."""
    n = len(jobs)
    scs = (C.c_void_p * n)(*[j[0].h.value for j in jobs])
    cams = (abi.Camera * n)(*[j[1] for j in jobs])
    ps = (abi.RenderParams * n)(*[j[2] for j in jobs])
    outs = [np.zeros((j[2].height, j[2].width, 5), np.float32) for j in jobs]
    ptrs = (abi.f32p * n)(*[abi.fptr(o) for o in outs])
    st = abi.Stats()
    check(lib().sgr_render_batch(ctx.h, n, scs, cams, ps, ptrs, 0, C.byref(st)),
          "sgr_render_batch")
    return outs, st.as_dict()


def octree_build(desc):
    """Host-only octree build: (nodes (N,4) u32, surf ids, origin+size, stats)."""
    nn, ns = C.c_uint64(), C.c_uint64()
    check(lib().sgr_octree_build(C.byref(desc.c()), None, C.byref(nn), None,
                                 C.byref(ns), None, None), "sgr_octree_build")
    nodes = np.zeros((nn.value, 4), np.uint32)
    surfs = np.zeros(ns.value, np.uint32)
    osz = np.zeros(4, np.float32)
    st = abi.OctreeStats()
    check(lib().sgr_octree_build(C.byref(desc.c()), abi.uptr(nodes), C.byref(nn),
                                 abi.uptr(surfs), C.byref(ns), abi.fptr(osz),
                                 C.byref(st)), "sgr_octree_build")
    return nodes, surfs, osz, {k: getattr(st, k) for k, _ in st._fields_}


def finalize(accum, exposure=1.0, contrast=1.0):
    a = np.ascontiguousarray(accum, np.float32)
    h, w = a.shape[:2]
    out = np.zeros((h, w, 4), np.float32)
    check(lib().sgr_finalize_image(abi.fptr(a), w, h, exposure, contrast, abi.fptr(out)),
          "sgr_finalize_image")
    return out


def _fmt(fmt):
    return None if fmt is None else fmt.encode()


def write_image(path, rgba, fmt=None):
    """Write a finished (H, W, 4) RGBA image as PFM or Radiance RGBE
    (format from the extension unless FMT is given)."""
    a = np.ascontiguousarray(rgba, np.float32)
    h, w = a.shape[:2]
    check(lib().sgr_image_write(str(path).encode(), _fmt(fmt), abi.fptr(a), w, h),
          "sgr_image_write")


def read_image(path, fmt=None):
    """(H, W, 3) float32 RGB of a PFM or RGBE file."""
    w, h = C.c_uint32(), C.c_uint32()
    check(lib().sgr_image_read(str(path).encode(), _fmt(fmt), None, C.byref(w), C.byref(h)),
          "sgr_image_read")
    out = np.zeros((h.value, w.value, 3), np.float32)
    check(lib().sgr_image_read(str(path).encode(), _fmt(fmt), abi.fptr(out), C.byref(w),
                               C.byref(h)), "sgr_image_read")
    return out


def recover_image(path, accum, fmt=None):
    """Fill the leading complete rows of a partially written output image
    into ACCUM ((H, W, 5) float32, modified in place); returns the number
    of rows recovered (a render of the rest continues at that row)."""
    assert accum.dtype == np.float32 and accum.flags.c_contiguous
    rows = C.c_uint32()
    h, w = accum.shape[:2]
    check(lib().sgr_image_recover(str(path).encode(), _fmt(fmt), abi.fptr(accum), w, h,
                                  C.byref(rows)), "sgr_image_recover")
    return rows.value


def cylinder_xform(base, axis, radius):
    """3x4 local-to-world transform of the cylinder with BASE, AXIS and
    RADIUS (Cylinder::xform, surface/cylinder.cc:24-36)."""
    b = np.ascontiguousarray(base, np.float32)
    a = np.ascontiguousarray(axis, np.float32)
    out = np.zeros(12, np.float32)
    check(lib().sgr_cylinder_xform(abi.fptr(b), abi.fptr(a), float(radius), abi.fptr(out)),
          "sgr_cylinder_xform")
    return out.reshape(3, 4)


def load_mesh(path, fmt=None):
    """(vertices (V, 3) float32, triangles (T, 3) uint32, normals (V, 3)
    float32 or None) of a PLY, OBJ, STL, MSH or UG file."""
    m = load_mesh_file(path, fmt)
    return m["v"], m["t"], m["n"]


def load_mesh_file(path, fmt=None):
    """A mesh file as a dict: v (V, 3) float32, t (T, 3) uint32, n (V, 3)
    float32 or None, left_handed (as the reference's loader leaves the
    mesh), tri_parts (T,) uint32 and part_colors (P, 3) float32 (a UG
    material's lambert colour, or -1 for the caller's material)."""
    h = C.c_void_p()
    nv, nt, hn = C.c_uint64(), C.c_uint64(), C.c_int()
    check(lib().sgr_mesh_file_open(str(path).encode(), _fmt(fmt), C.byref(h), C.byref(nv),
                                   C.byref(nt), C.byref(hn)), "sgr_mesh_file_open")
    try:
        v = np.zeros((nv.value, 3), np.float32)
        t = np.zeros((nt.value, 3), np.uint32)
        n = np.zeros((nv.value, 3), np.float32) if hn.value else None
        check(lib().sgr_mesh_file_copy(h, abi.fptr(v), abi.uptr(t),
                                       abi.fptr(n) if n is not None else None),
              "sgr_mesh_file_copy")
        lh, npart = C.c_int(), C.c_uint32()
        check(lib().sgr_mesh_file_parts(h, C.byref(lh), C.byref(npart), None, None),
              "sgr_mesh_file_parts")
        tp = np.zeros(nt.value, np.uint32)
        pc = np.zeros((npart.value, 3), np.float32)
        check(lib().sgr_mesh_file_parts(h, None, None, abi.uptr(tp), abi.fptr(pc)),
              "sgr_mesh_file_parts")
    finally:
        lib().sgr_mesh_file_close(h)
    return dict(v=v, t=t, n=n, left_handed=bool(lh.value), tri_parts=tp, part_colors=pc)
