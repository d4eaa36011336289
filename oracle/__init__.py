"""ORACLE -- test infrastructure only.

ctypes front end of liboracle.so, the CPU restatement of the reference's
path-tracing render path.  Only tests/, the smoke check and the CPU
baseline leg of bench.py may import this package; the product never does.
"""
import ctypes as C
import os
import subprocess

import numpy as np

from snogray_amd import abi

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
# the shipped-flags build with the negative mailbox: CPU-baseline timing only
LIB_FAST = os.path.join(HERE, "liboracle_fast.so")

_libs = {}


def build():
    subprocess.check_call(["make", "-s", "-C", HERE])


def lib(fast=False):
    if fast not in _libs:
        path = LIB_FAST if fast else LIB
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        L.orc_scene_create.restype = C.c_void_p
        L.orc_scene_create.argtypes = [C.POINTER(abi.SceneDescC), C.c_int,
                                       C.c_char_p, C.c_int]
        L.orc_scene_destroy.argtypes = [C.c_void_p]
        L.orc_octree_export.argtypes = [C.c_void_p, abi.u32p,
                                        C.POINTER(C.c_uint64), abi.u32p,
                                        C.POINTER(C.c_uint64),
                                        C.POINTER(C.c_double)]
        L.orc_octree_stats.argtypes = [C.c_void_p, C.POINTER(abi.OctreeStats)]
        L.orc_scene_bounds.argtypes = [C.c_void_p] + [C.POINTER(C.c_double)] * 3
        L.orc_trace.argtypes = [C.c_void_p, C.POINTER(abi.Rays), C.c_size_t,
                                C.POINTER(abi.Hits), C.POINTER(C.c_uint8),
                                C.c_int, C.POINTER(C.c_uint64)]
        L.orc_render.argtypes = [C.c_void_p, C.POINTER(abi.Camera),
                                 C.POINTER(abi.RenderParams),
                                 C.POINTER(C.c_double), abi.f32p, C.c_int,
                                 C.POINTER(abi.Stats), C.c_char_p, C.c_int]
        L.orc_bsdf_probe.argtypes = [C.POINTER(abi.Material)] + \
            [C.POINTER(C.c_double)] * 2 + [C.c_int, C.c_double, C.c_double,
                                           C.POINTER(C.c_double), C.c_int,
                                           C.POINTER(C.c_double),
                                           C.POINTER(C.c_double), C.c_int,
                                           C.POINTER(C.c_double), C.c_int,
                                           C.c_int]
        L.orc_fresnel.restype = C.c_double
        L.orc_fresnel.argtypes = [C.c_double] * 4
        L.orc_mitchell.restype = C.c_double
        L.orc_mitchell.argtypes = [C.c_double] * 4
        L.orc_envmap_probe.argtypes = [abi.f32p, C.c_uint32, C.c_uint32,
                                       C.POINTER(C.c_double), C.c_int,
                                       C.POINTER(C.c_double)]
        L.orc_envmap_lookup.argtypes = [abi.f32p, C.c_uint32, C.c_uint32,
                                        C.POINTER(C.c_double), C.c_int,
                                        C.POINTER(C.c_double)]
        L.orc_stream_probe.argtypes = [C.c_uint64, C.c_int, C.c_int, C.c_uint32,
                                       C.c_uint32, abi.f32p, C.c_uint32,
                                       C.c_uint32, abi.f32p]
        L.orc_isec_probe.argtypes = [C.c_void_p, abi.f32p, C.c_int, C.POINTER(C.c_double)]
        L.orc_light_probe.argtypes = [C.c_void_p, C.c_uint, C.c_int, C.POINTER(C.c_double),
                                      C.c_int, C.POINTER(C.c_double)]
        L.orc_photon_map_create.restype = C.c_void_p
        L.orc_photon_map_create.argtypes = [C.c_void_p, C.POINTER(abi.PhotonParams), C.c_char_p,
                                            C.c_int]
        L.orc_photon_map_import.restype = C.c_void_p
        L.orc_photon_map_import.argtypes = [C.POINTER(abi.PhotonParams)] + \
            [C.POINTER(abi.f32p)] * 3 + [C.POINTER(C.c_uint64)] * 2
        L.orc_photon_map_info.argtypes = [C.c_void_p, C.POINTER(abi.PhotonMapInfo)]
        L.orc_photon_map_export.argtypes = [C.c_void_p, C.c_int, abi.f32p, abi.f32p, abi.f32p]
        L.orc_photon_map_destroy.argtypes = [C.c_void_p]
        L.orc_radical_inverse.restype = C.c_double
        L.orc_radical_inverse.argtypes = [C.c_uint, C.c_uint]
        L.orc_kd_probe.argtypes = [abi.f32p, C.c_int, abi.f32p, C.c_int, C.c_uint, C.c_float,
                                   abi.f32p, C.POINTER(C.c_int), C.POINTER(C.c_int), abi.f32p]
        L.orc_dirhist_probe.argtypes = [abi.f32p, abi.f32p, C.c_int, abi.f32p, C.c_int,
                                        abi.f32p, abi.f32p]
        L.orc_permute.restype = C.c_uint32
        L.orc_permute.argtypes = [C.c_uint32, C.c_uint32, C.c_uint64]
        _libs[fast] = L
    return _libs[fast]


def dptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class OracleScene:
    """CPU-oracle scene (fp32 reproduces the reference arithmetic, fp64 is
    the high-precision oracle).  fast: the shipped-flags build with the
    negative mailbox, for timing only (same images up to fast-math
    rounding; surface-test counts then exclude the mailbox's skips)."""

    def __init__(self, desc, fp64=False, fast=False):
        err = C.create_string_buffer(512)
        self.desc = desc
        self.L = lib(fast)
        self.h = self.L.orc_scene_create(C.byref(desc.c()), 1 if fp64 else 0,
                                         err, 512)
        if not self.h:
            raise RuntimeError("oracle scene: " + err.value.decode())
        self.fp64 = fp64

    def __del__(self):
        if getattr(self, "h", None):
            self.L.orc_scene_destroy(self.h)
            self.h = None

    def octree(self):
        nn, ns = C.c_uint64(), C.c_uint64()
        self.L.orc_octree_export(self.h, None, C.byref(nn), None, C.byref(ns), None)
        nodes = np.zeros((nn.value, 4), np.uint32)
        surfs = np.zeros(ns.value, np.uint32)
        osz = np.zeros(4, np.float64)
        self.L.orc_octree_export(self.h, abi.uptr(nodes), C.byref(nn),
                                abi.uptr(surfs), C.byref(ns), dptr(osz))
        return nodes, surfs, osz

    def octree_stats(self):
        st = abi.OctreeStats()
        self.L.orc_octree_stats(self.h, C.byref(st))
        return {k: getattr(st, k) for k, _ in st._fields_}

    def bounds(self):
        a, b, h = np.zeros(3), np.zeros(3), C.c_double()
        self.L.orc_scene_bounds(self.h, dptr(a), dptr(b), C.byref(h))
        return a, b, h.value

    def trace_closest(self, batch, nthreads=8):
        hb = abi.HitBuffer(batch.n)
        cnt = (C.c_uint64 * 2)()
        self.L.orc_trace(self.h, C.byref(batch.c), batch.n, C.byref(hb.c), None,
                        nthreads, cnt)
        return hb, (cnt[0], cnt[1])

    def trace_occluded(self, batch, nthreads=8):
        occ = np.zeros(batch.n, np.uint8)
        cnt = (C.c_uint64 * 2)()
        self.L.orc_trace(self.h, C.byref(batch.c), batch.n, None,
                        occ.ctypes.data_as(C.POINTER(C.c_uint8)), nthreads, cnt)
        return occ, (cnt[0], cnt[1])

    def isec_probe(self, origins, dirs, t0=1e-3, t1=1e9):
        """Closest hit and hit frame of each ray (fp32 scene): dict of
        prim, inst (-1: none), t, origin, normal (frame z), back, u, v."""
        o = np.asarray(origins, np.float32).reshape(-1, 3)
        d = np.asarray(dirs, np.float32).reshape(-1, 3)
        rays = np.zeros((len(o), 8), np.float32)
        rays[:, :3], rays[:, 3:6], rays[:, 6], rays[:, 7] = o, d, t0, t1
        out = np.zeros((len(o), 12))
        if self.L.orc_isec_probe(self.h, abi.fptr(rays), len(o), dptr(out)):
            raise RuntimeError("isec probe needs an fp32 oracle scene")
        return {"prim": out[:, 0].astype(np.int64), "inst": out[:, 1].astype(np.int64),
                "t": out[:, 2], "origin": out[:, 3:6], "normal": out[:, 6:9],
                "back": out[:, 9] > 0, "u": out[:, 10], "v": out[:, 11]}

    def light_probe(self, light, points, normals, uv=None, dirs=None):
        """Light LIGHT from POINTS with NORMALS: Light::Sampler::sample at
        parameters UV (n x 2), or, with DIRS (world directions, n x 3),
        Light::Sampler::eval.  Returns a dict of val (n x 3), pdf, dist and
        (sampling) the world direction."""
        pts = np.asarray(points, np.float64).reshape(-1, 3)
        n = len(pts)
        q = np.zeros((n, 8))
        q[:, 0:3] = pts
        q[:, 3:6] = np.asarray(normals, np.float64).reshape(-1, 3)
        if uv is not None:
            q[:, 6:8] = np.asarray(uv, np.float64).reshape(-1, 2)
        mode = 0 if dirs is None else 1
        if mode:
            q = np.concatenate([q.ravel(), np.asarray(dirs, np.float64).reshape(-1, 3).ravel()])
        q = np.ascontiguousarray(q.ravel())
        out = np.zeros((n, 8))
        if self.L.orc_light_probe(self.h, light, mode, dptr(q), n, dptr(out)):
            raise RuntimeError("light probe: no such light")
        return {"val": out[:, 0:3], "pdf": out[:, 3], "dist": out[:, 4], "dir": out[:, 5:8]}

    def photon_map(self, params=None):
        """The oracle's own photon maps of this scene."""
        return OraclePhotonMap(self, params if params is not None else abi.photon_params())

    def render(self, cam, params, nthreads=8, samples=False):
        acc = np.zeros((params.height, params.width, 5), np.float64)
        x0, y0, w, h, m = abi.render_window(params)
        out = None
        if samples:
            out = np.zeros(((h + 2 * m) * (w + 2 * m) * params.spp, 6), np.float32)
        st = abi.Stats()
        err = C.create_string_buffer(512)
        rc = self.L.orc_render(self.h, C.byref(cam), C.byref(params), dptr(acc),
                              abi.fptr(out) if out is not None else None,
                              nthreads, C.byref(st), err, 512)
        if rc:
            raise RuntimeError("oracle render: " + err.value.decode())
        return acc, out, st.as_dict()


def bsdf_probe(mat, v, gn, params, dirs, back=False, media_ior=1.0,
               enclosing_ior=1.0, flags_s=0x37, flags_e=0x37):
    m = abi.Material()
    m.type = mat["type"]
    m.color[:] = list(mat.get("color", (0, 0, 0)))
    m.gloss_color[:] = list(mat.get("gloss_color", (0, 0, 0)))
    m.m, m.ior_n, m.ior_k = mat.get("m", 0), mat.get("ior_n", 1.5), mat.get("ior_k", 0)
    m.underlying = -1
    v = np.ascontiguousarray(v, np.float64)
    gn = np.ascontiguousarray(gn, np.float64)
    params = np.ascontiguousarray(params, np.float64).reshape(-1, 2)
    dirs = np.ascontiguousarray(dirs, np.float64).reshape(-1, 3)
    out_s = np.zeros((len(params), 8))
    out_e = np.zeros((len(dirs), 4))
    lib().orc_bsdf_probe(C.byref(m), dptr(v), dptr(gn), 1 if back else 0,
                         media_ior, enclosing_ior, dptr(params), len(params),
                         dptr(out_s), dptr(dirs), len(dirs), dptr(out_e),
                         flags_s, flags_e)
    return out_s, out_e


def fresnel(n1, n2, k2, cos_angle):
    return lib().orc_fresnel(n1, n2, k2, cos_angle)


def mitchell(x, radius=2.0, b=1.0 / 3, c=1.0 / 3):
    return lib().orc_mitchell(x, radius, b, c)


def envmap_probe(texels, params):
    t = np.ascontiguousarray(texels, np.float32)
    p = np.ascontiguousarray(params, np.float64).reshape(-1, 2)
    out = np.zeros((len(p), 10))
    lib().orc_envmap_probe(abi.fptr(t), t.shape[1], t.shape[0], dptr(p), len(p),
                           dptr(out))
    return out


def envmap_lookup(texels, uv):
    """Bilinear lookup of an (H, W, 3) map (row 0 = top) at (u, v) pairs."""
    t = np.ascontiguousarray(texels, np.float32)
    p = np.ascontiguousarray(uv, np.float64).reshape(-1, 2)
    out = np.zeros((len(p), 3))
    lib().orc_envmap_lookup(abi.fptr(t), t.shape[1], t.shape[0], dptr(p), len(p), dptr(out))
    return out


def stream_probe(seed, px, py, chan, spp, s=0, nr=0):
    uv = np.zeros((spp, 2), np.float32)
    r = np.zeros(max(nr, 1), np.float32)
    lib().orc_stream_probe(seed, px, py, chan, spp, abi.fptr(uv), s, nr,
                           abi.fptr(r))
    return uv, r[:nr]


def permute(i, n, key):
    return lib().orc_permute(i, n, key)


class OraclePhotonMap:
    """Photon maps shot and built by the oracle, or (import_arrays) taken
    from another builder's kd-ordered export; pass it to
    abi.make_params(photon_map=...) for an oracle render."""

    def __init__(self, scene=None, params=None, arrays=None, paths=None):
        self.L = lib()
        self.params = params
        err = C.create_string_buffer(512)
        if scene is not None:
            self.h = self.L.orc_photon_map_create(scene.h, C.byref(params), err, 512)
            if not self.h:
                raise RuntimeError(err.value.decode())
        else:
            self._keep = [[np.ascontiguousarray(a, np.float32) for a in st] for st in arrays]
            pa = (abi.f32p * 3)(*[abi.fptr(st[0]) for st in self._keep])
            pw = (abi.f32p * 3)(*[abi.fptr(st[1]) for st in self._keep])
            dr = (abi.f32p * 3)(*[abi.fptr(st[2]) for st in self._keep])
            cnt = (C.c_uint64 * 3)(*[len(st[0]) for st in self._keep])
            pth = (C.c_uint64 * 3)(*[int(x) for x in paths])
            self.h = self.L.orc_photon_map_import(C.byref(params), pa, pw, dr, cnt, pth)

    @classmethod
    def import_arrays(cls, params, arrays, paths):
        """ARRAYS: per set (pos_axis (n, 4), power (n, 3), dir (n, 3)) in
        kd-tree order; PATHS: each set's path count."""
        return cls(None, params, arrays, paths)

    def handle(self):
        return self.h

    def info(self):
        i = abi.PhotonMapInfo()
        self.L.orc_photon_map_info(self.h, C.byref(i))
        return i.as_dict()

    def export(self, which):
        n = self.info()["photons"][which]
        pa = np.zeros((n, 4), np.float32)
        pw = np.zeros((n, 3), np.float32)
        dr = np.zeros((n, 3), np.float32)
        self.L.orc_photon_map_export(self.h, which, abi.fptr(pa), abi.fptr(pw), abi.fptr(dr))
        return pa, pw, dr

    def __del__(self):
        if getattr(self, "h", None):
            self.L.orc_photon_map_destroy(self.h)
            self.h = None


def radical_inverse(num, base):
    return lib().orc_radical_inverse(num, base)


def kd_probe(points, queries, k, r2):
    """The oracle's photon kd-tree over POINTS and its k-nearest search
    from QUERIES: (heap-order positions, split axes, found kd indices
    (m, k; -1 padded), final squared radii)."""
    p = np.ascontiguousarray(points, np.float32).reshape(-1, 3)
    q = np.ascontiguousarray(queries, np.float32).reshape(-1, 3)
    kp = np.zeros_like(p)
    ax = np.zeros(len(p), np.int32)
    found = np.zeros((len(q), k), np.int32)
    r = np.zeros(len(q), np.float32)
    ip = lambda a: a.ctypes.data_as(C.POINTER(C.c_int))
    lib().orc_kd_probe(abi.fptr(p), len(p), abi.fptr(q), len(q), k, r2, abi.fptr(kp), ip(ax),
                       ip(found), abi.fptr(r))
    return kp, ax, found, r


def dirhist_probe(dirs, weights, params):
    """Direction histogram of DIRS with WEIGHTS: samples at PARAMS (n, 5:
    direction, pdf, pdf() at it) and the 64 bin pdfs."""
    d = np.ascontiguousarray(dirs, np.float32).reshape(-1, 3)
    w = np.ascontiguousarray(weights, np.float32)
    pr = np.ascontiguousarray(params, np.float32).reshape(-1, 2)
    out = np.zeros((len(pr), 5), np.float32)
    bins = np.zeros(64, np.float32)
    lib().orc_dirhist_probe(abi.fptr(d), abi.fptr(w), len(d), abi.fptr(pr), len(pr),
                            abi.fptr(out), abi.fptr(bins))
    return out, bins
