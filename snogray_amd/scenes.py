"""Scene descriptions and the seeded synthetic scenes.

``SceneDesc`` collects meshes, tripars, spheres, materials and envmaps as
numpy arrays and produces the C descriptor of include/snoggpu.h; the
same object feeds the device library and the CPU oracle.

Synthetic scenes (all sizes in metres, y is up):

* ``cornell_box``   -- 5.5 m Cornell box: 10 Lambert wall triangles,
  two Lambert boxes of 12 triangles, a 1.3 x 1.05 m parallelogram
  ceiling light Glow(17, 12, 4).
* ``icosphere_soup`` -- N level-L icospheres (radius 1 m) radially
  displaced by seeded 3-octave value noise (amplitude 5 %), centres
  uniform in a 20 m cube above a 2-triangle ground; 80 % Cook-Torrance
  (m ~ U[0.05, 0.5], IOR ~ U[1.3, 1.8]), 20 % glass IOR 1.5; lit by a
  Gaussian-lobe lat-long envmap (0.1 sky + 6 lobes, peaks <= 1e3).
* ``furnace_box``   -- closed box of emitting Lambert parallelograms
  (closed-form radiance, used to pin the integrator).
"""
import ctypes as C
import math

import numpy as np

from . import abi

Y_UP_ENV_FRAME = (1.0, 0.0, 0.0,   # map x -> world x
                  0.0, 0.0, 1.0,   # map y -> world z
                  0.0, 1.0, 0.0)   # map z (north) -> world y


class SceneDesc:
    """Host-side scene description (keeps every array alive)."""

    def __init__(self, min_trace=1e-3, accel=abi.SGR_ACCEL_OCTREE):
        self.materials = []
        self.meshes = []
        self.tripars = []
        self.spheres = []
        self.ellipses = []
        self.local_prims = []  # (SGR_LOCAL_* kind, 3x4 local-to-world, material)
        self.envmaps = []
        self.point_lights = []
        self.models = []       # SceneDesc objects holding geometry only
        self.instances = []    # (model index, 3x4 local-to-world)
        self.min_trace = min_trace
        self.accel = accel
        self._c = None

    # --- materials (material/*.cc) -----------------------------------
    def _mat(self, **kw):
        m = dict(type=0, color=(0, 0, 0), gloss_color=(0, 0, 0), m=0.0,
                 ior_n=1.5, ior_k=0.0, underlying=-1)
        m.update(kw)
        self.materials.append(m)
        self._c = None
        return len(self.materials) - 1

    def lambert(self, color):
        return self._mat(type=abi.SGR_MAT_LAMBERT, color=color)

    def cook_torrance(self, color, gloss_color, m, ior, ior_k=0.0):
        return self._mat(type=abi.SGR_MAT_COOK_TORRANCE, color=color,
                         gloss_color=gloss_color, m=m, ior_n=ior, ior_k=ior_k)

    def glass(self, ior):
        return self._mat(type=abi.SGR_MAT_GLASS, ior_n=ior)

    def mirror(self, reflectance, ior=1.5, ior_k=0.0, color=(0, 0, 0), underlying=-1):
        """Mirror (material/mirror.h): over a Lambert of COLOR, or over the
        material UNDERLYING (a Lambert, Cook-Torrance or Phong)."""
        return self._mat(type=abi.SGR_MAT_MIRROR, gloss_color=reflectance,
                         ior_n=ior, ior_k=ior_k, color=color, underlying=underlying)

    def phong(self, color, specular_color, exponent):
        return self._mat(type=abi.SGR_MAT_PHONG, color=color,
                         gloss_color=specular_color, m=exponent)

    def thin_glass(self, color=(1.0, 1.0, 1.0), ior=1.5):
        return self._mat(type=abi.SGR_MAT_THIN_GLASS, color=color, ior_n=ior)

    def stencil(self, opacity, underlying):
        return self._mat(type=abi.SGR_MAT_STENCIL, color=opacity, underlying=underlying)

    def glow(self, color, underlying=-1):
        return self._mat(type=abi.SGR_MAT_GLOW, color=color,
                         underlying=underlying)

    # --- surfaces ------------------------------------------------------
    def mesh(self, vertices, triangles, material, tri_materials=None,
             axis=(0.0, 0.0, 1.0), left_handed=True, normals=None):
        v = np.ascontiguousarray(vertices, dtype=np.float32).reshape(-1, 3)
        t = np.ascontiguousarray(triangles, dtype=np.uint32).reshape(-1, 3)
        tm = None if tri_materials is None else np.ascontiguousarray(
            tri_materials, dtype=np.uint32)
        vn = None if normals is None else np.ascontiguousarray(
            normals, dtype=np.float32).reshape(-1, 3)
        assert vn is None or len(vn) == len(v), "one normal per vertex"
        self.meshes.append(dict(v=v, t=t, tm=tm, material=material,
                                axis=axis, left_handed=left_handed, vn=vn))
        self._c = None
        return len(self.meshes) - 1

    def mesh_file(self, path, material, axis=(0.0, 0.0, 1.0), fmt=None):
        """A PLY, OBJ, STL, MSH or UG mesh file (with its vertex normals, if
        any) as one mesh, handed as the reference's loader leaves it
        (load/load.lua:405-418).  A UG file's named colours become lambert
        materials, one per part (load/mesh/ug.lua:55-90); every other
        triangle takes MATERIAL."""
        from . import native
        f = native.load_mesh_file(path, fmt)
        tm = None
        if len(f["part_colors"]) > 1 or (len(f["part_colors"]) and f["part_colors"][0][0] >= 0):
            mats = np.array([material if c[0] < 0 else self.lambert(tuple(map(float, c)))
                             for c in f["part_colors"]], np.uint32)
            tm = mats[f["tri_parts"]]
        return self.mesh(f["v"], f["t"], material, tri_materials=tm, axis=axis,
                         left_handed=f["left_handed"], normals=f["n"])

    def model(self):
        """A new model: add geometry to the returned SceneDesc (its
        material indices are this scene's), then place it with instance()."""
        m = SceneDesc(min_trace=self.min_trace, accel=self.accel)
        self.models.append(m)
        self._c = None
        return len(self.models) - 1, m

    def instance(self, model, local_to_world):
        """Place MODEL by a 3x4 (or 4x4) affine local-to-world transform."""
        x = np.asarray(local_to_world, np.float64)[:3, :4]
        self.instances.append((int(model), x.astype(np.float32)))
        self._c = None

    def tripar(self, v0, e1, e2, material, parallelogram=False):
        self.tripars.append(dict(v0=v0, e1=e1, e2=e2, material=material,
                                 parallelogram=parallelogram))
        self._c = None

    def sphere(self, center, radius, material):
        self.spheres.append(dict(center=center, radius=radius,
                                 material=material))
        self._c = None

    def ellipse(self, center, radius1, radius2, material):
        """Ellipse (surface/ellipse.h:38-45) at CENTER with radius vectors
        RADIUS1 and RADIUS2."""
        self.ellipses.append(dict(center=center, r1=radius1, r2=radius2,
                                  material=material))
        self._c = None

    def _local(self, kind, local_to_world, material):
        x = np.asarray(local_to_world, np.float64)[:3, :4].astype(np.float32)
        self.local_prims.append((kind, x, int(material)))
        self._c = None

    def cylinder(self, local_to_world, material):
        """Cylinder (surface/cylinder.h:24-33): the canonical cylinder
        (radius 1 about z, z in [-1, 1], no caps) placed by a 3x4 transform;
        cylinder_xform gives the transform of a base, axis and radius."""
        self._local(abi.SGR_LOCAL_CYLINDER, local_to_world, material)

    def sphere2(self, local_to_world, material):
        """Sphere2 (surface/sphere2.h:24-33): the unit sphere placed by a
        3x4 transform."""
        self._local(abi.SGR_LOCAL_SPHERE2, local_to_world, material)

    def point_light(self, pos, intensity, angle=0.0, dir=(0.0, 0.0, 1.0), fringe_angle=0.0):
        """Point light (light/point-light.h:31-47, light.point in
        light/light.lua:48-56): INTENSITY from POS in all directions, or
        with ANGLE > 0 a spotlight of that apex angle about DIR whose outer
        FRINGE_ANGLE falls off toward the edge."""
        self.point_lights.append(dict(pos=pos, intensity=intensity, angle=float(angle),
                                      dir=dir, fringe_angle=float(fringe_angle)))
        self._c = None

    def envmap(self, texels, frame=Y_UP_ENV_FRAME):
        t = np.ascontiguousarray(texels, dtype=np.float32)
        assert t.ndim == 3 and t.shape[2] == 3
        self.envmaps.append(dict(tex=t, frame=frame))
        self._c = None

    # --- counts --------------------------------------------------------
    @property
    def num_triangles(self):
        return sum(len(m["t"]) for m in self.meshes)

    @property
    def num_prims(self):
        return (self.num_triangles + len(self.tripars) + len(self.ellipses) +
                len(self.spheres) + len(self.local_prims))

    # --- C descriptor ----------------------------------------------------
    def _geometry(self):
        """C arrays of this description's meshes, tripars and spheres."""
        meshes = (abi.Mesh * max(1, len(self.meshes)))()
        for i, m in enumerate(self.meshes):
            meshes[i].vertices = abi.fptr(m["v"])
            meshes[i].num_vertices = len(m["v"])
            meshes[i].triangles = abi.uptr(m["t"])
            meshes[i].num_triangles = len(m["t"])
            meshes[i].tri_materials = (abi.uptr(m["tm"]) if m["tm"] is not None
                                       else C.cast(None, abi.u32p))
            meshes[i].material = m["material"]
            meshes[i].axis[:] = [float(x) for x in m["axis"]]
            meshes[i].left_handed = 1 if m["left_handed"] else 0
            meshes[i].normals = abi.fptr(m["vn"]) if m.get("vn") is not None else None
        tps = (abi.Tripar * max(1, len(self.tripars)))()
        for i, t in enumerate(self.tripars):
            tps[i].v0[:] = [float(x) for x in t["v0"]]
            tps[i].e1[:] = [float(x) for x in t["e1"]]
            tps[i].e2[:] = [float(x) for x in t["e2"]]
            tps[i].parallelogram = 1 if t["parallelogram"] else 0
            tps[i].material = t["material"]
        sps = (abi.Sphere * max(1, len(self.spheres)))()
        for i, s in enumerate(self.spheres):
            sps[i].center[:] = [float(x) for x in s["center"]]
            sps[i].radius = s["radius"]
            sps[i].material = s["material"]
        els = (abi.Ellipse * max(1, len(self.ellipses)))()
        for i, e in enumerate(self.ellipses):
            els[i].center[:] = [float(x) for x in e["center"]]
            els[i].radius1[:] = [float(x) for x in e["r1"]]
            els[i].radius2[:] = [float(x) for x in e["r2"]]
            els[i].material = e["material"]
        lps = (abi.LocalPrim * max(1, len(self.local_prims)))()
        for i, (kind, x, mat) in enumerate(self.local_prims):
            lps[i].kind = kind
            lps[i].local_to_world[:] = [float(v) for v in x.ravel()]
            lps[i].material = mat
        return meshes, tps, sps, els, lps

    def c(self):
        if self._c is not None:
            return self._c
        mats = (abi.Material * max(1, len(self.materials)))()
        for i, m in enumerate(self.materials):
            mats[i].type = m["type"]
            mats[i].color[:] = [float(x) for x in m["color"]]
            mats[i].gloss_color[:] = [float(x) for x in m["gloss_color"]]
            mats[i].m, mats[i].ior_n, mats[i].ior_k = m["m"], m["ior_n"], m["ior_k"]
            mats[i].underlying = m["underlying"]
        meshes, tps, sps, els, lps = self._geometry()
        envs = (abi.Envmap * max(1, len(self.envmaps)))()
        for i, e in enumerate(self.envmaps):
            envs[i].texels = abi.fptr(e["tex"])
            envs[i].height, envs[i].width = e["tex"].shape[0], e["tex"].shape[1]
            envs[i].frame[:] = [float(x) for x in e["frame"]]
        pls = (abi.PointLight * max(1, len(self.point_lights)))()
        for i, pl in enumerate(self.point_lights):
            pls[i].pos[:] = [float(x) for x in pl["pos"]]
            pls[i].intensity[:] = [float(x) for x in pl["intensity"]]
            pls[i].dir[:] = [float(x) for x in pl["dir"]]
            pls[i].angle, pls[i].fringe_angle = pl["angle"], pl["fringe_angle"]
        d = abi.SceneDescC()
        d.meshes, d.num_meshes = meshes, len(self.meshes)
        d.tripars, d.num_tripars = tps, len(self.tripars)
        d.spheres, d.num_spheres = sps, len(self.spheres)
        d.ellipses, d.num_ellipses = els, len(self.ellipses)
        d.local_prims, d.num_local_prims = lps, len(self.local_prims)
        d.materials, d.num_materials = mats, len(self.materials)
        d.envmaps, d.num_envmaps = envs, len(self.envmaps)
        d.point_lights, d.num_point_lights = pls, len(self.point_lights)
        d.min_trace, d.accel = self.min_trace, self.accel
        models = (abi.Model * max(1, len(self.models)))()
        keep_models = []
        for i, m in enumerate(self.models):
            g = m._geometry()
            keep_models.append(g)
            models[i].meshes, models[i].num_meshes = g[0], len(m.meshes)
            models[i].tripars, models[i].num_tripars = g[1], len(m.tripars)
            models[i].spheres, models[i].num_spheres = g[2], len(m.spheres)
            models[i].ellipses, models[i].num_ellipses = g[3], len(m.ellipses)
            models[i].local_prims, models[i].num_local_prims = g[4], len(m.local_prims)
        insts = (abi.Instance * max(1, len(self.instances)))()
        for i, (mi, x) in enumerate(self.instances):
            insts[i].model = mi
            insts[i].local_to_world[:] = [float(v) for v in x.ravel()]
        d.models = C.cast(models, C.c_void_p) if self.models else None
        d.num_models = len(self.models)
        d.instances = C.cast(insts, C.c_void_p) if self.instances else None
        d.num_instances = len(self.instances)
        self._keep = (mats, meshes, tps, sps, els, lps, envs, pls, models, keep_models, insts)
        self._c = d
        return d

    # --- geometry access for tests ----------------------------------------
    def prim_table(self):
        """(kind, v0, e1, e2, center, radius) arrays indexed by prim id,
        with e1/e2 computed in float32 as the intersection code does."""
        v0s, e1s, e2s = [], [], []
        for m in self.meshes:
            v, t = m["v"], m["t"]
            a, b, c = v[t[:, 0]], v[t[:, 1]], v[t[:, 2]]
            v0s.append(a)
            e1s.append((b - a).astype(np.float32))
            e2s.append((c - a).astype(np.float32))
        for t in self.tripars:
            v0s.append(np.array([t["v0"]], np.float32))
            e1s.append(np.array([t["e1"]], np.float32))
            e2s.append(np.array([t["e2"]], np.float32))
        if v0s:
            return (np.concatenate(v0s), np.concatenate(e1s),
                    np.concatenate(e2s))
        z = np.zeros((0, 3), np.float32)
        return z, z, z

    def bbox(self):
        pts = [m["v"] for m in self.meshes]
        for t in self.tripars:
            v0 = np.array(t["v0"], np.float32)
            e1 = np.array(t["e1"], np.float32)
            e2 = np.array(t["e2"], np.float32)
            pts.append(np.stack([v0, v0 + e1, v0 + e2, v0 + e1 + e2]))
        for s in self.spheres:
            c, r = np.array(s["center"], np.float32), s["radius"]
            pts.append(np.stack([c - r, c + r]))
        p = np.concatenate(pts)
        return p.min(0), p.max(0)


def camera(pos, forward, up, right, tan_half_fov_x, tan_half_fov_y):
    c = abi.Camera()
    c.pos[:] = [float(x) for x in pos]
    c.forward[:] = [float(x) for x in forward]
    c.up[:] = [float(x) for x in up]
    c.right[:] = [float(x) for x in right]
    c.tan_half_fov_x, c.tan_half_fov_y = tan_half_fov_x, tan_half_fov_y
    return c


def look_at(pos, target, up=(0.0, 1.0, 0.0), vfov_deg=40.0, aspect=1.0):
    """Camera at POS looking at TARGET (left-handed: right = up x forward,
    as snogray's default frame forward=+z, up=+y, right=+x)."""
    p = np.asarray(pos, np.float64)
    f = np.asarray(target, np.float64) - p
    f /= np.linalg.norm(f)
    u0 = np.asarray(up, np.float64)
    r = np.cross(u0, f)
    r /= np.linalg.norm(r)
    u = np.cross(f, r)
    t = math.tan(math.radians(vfov_deg) / 2)
    return camera(p, f, u, r, t * aspect, t)


# ---------------------------------------------------------------------------
# Cornell box
# ---------------------------------------------------------------------------

def _quad_tris(a, b, c, d):
    """Two triangles (a, b, c), (a, c, d)."""
    return [a, b, c], [a, c, d]


def _box_mesh(center_xz, size, height, angle_deg):
    """12-triangle box standing on y=0, rotated about y."""
    cx, cz = center_xz
    s = size / 2.0
    ang = math.radians(angle_deg)
    ca, sa = math.cos(ang), math.sin(ang)
    corners = []
    for y in (0.0, height):
        for (x, z) in ((-s, -s), (s, -s), (s, s), (-s, s)):
            corners.append((cx + ca * x - sa * z, y, cz + sa * x + ca * z))
    v = np.array(corners, np.float32)
    faces = [(0, 3, 2, 1), (4, 5, 6, 7), (0, 1, 5, 4), (1, 2, 6, 5),
             (2, 3, 7, 6), (3, 0, 4, 7)]
    tris = []
    for a, b, c, d in faces:
        tris += [[a, b, c], [a, c, d]]
    return v, np.array(tris, np.uint32)


def cornell_box(light_emission=(17.0, 12.0, 4.0), aspect=1.0):
    """Cornell box, 5.5 m, y up; returns (SceneDesc, camera)."""
    S = 5.5
    sd = SceneDesc()
    white = sd.lambert((0.73, 0.73, 0.73))
    red = sd.lambert((0.63, 0.065, 0.05))
    green = sd.lambert((0.14, 0.45, 0.09))
    light = sd.glow(light_emission)
    P = lambda x, y, z: (x, y, z)
    walls_v = np.array([
        P(0, 0, 0), P(S, 0, 0), P(S, 0, S), P(0, 0, S),      # floor 0-3
        P(0, S, 0), P(S, S, 0), P(S, S, S), P(0, S, S),      # ceiling 4-7
    ], np.float32)
    tris, mats = [], []
    for q, m in (((0, 1, 2, 3), white),     # floor
                 ((4, 7, 6, 5), white),     # ceiling
                 ((3, 2, 6, 7), white),     # back wall z = S
                 ((0, 3, 7, 4), red),       # left wall x = 0
                 ((1, 5, 6, 2), green)):    # right wall x = S
        a, b, c, d = q
        tris += [[a, b, c], [a, c, d]]
        mats += [m, m]
    sd.mesh(walls_v, np.array(tris, np.uint32), white,
            tri_materials=np.array(mats, np.uint32))
    v, t = _box_mesh((1.85, 1.69), 1.65, 1.65, -17.0)
    sd.mesh(v, t, white)
    v, t = _box_mesh((3.68, 3.51), 1.65, 3.30, 17.0)
    sd.mesh(v, t, white)
    # Ceiling light, facing down: normal = cross(e2, e1) = -y.
    lx, lz = 1.3, 1.05
    sd.tripar((S / 2 - lx / 2, S - 0.01, S / 2 - lz / 2), (0.0, 0.0, lz),
              (lx, 0.0, 0.0), light, parallelogram=True)
    t = math.tan(math.radians(39.3) / 2)
    cam = camera((S / 2, S / 2, -8.0), (0, 0, 1), (0, 1, 0), (1, 0, 0), t * aspect, t)
    return sd, cam


def furnace_box(albedo=0.5, emission=1.0, size=2.0):
    """Closed box of six emitting Lambert parallelograms facing inward.
    With every wall emitting E and reflecting albedo a, PathInteg's
    expected radiance at max_path_len M is E * sum_{j=0}^{M+1} a^j."""
    sd = SceneDesc()
    under = sd.lambert((albedo, albedo, albedo))
    g = sd.glow((emission, emission, emission), underlying=under)
    s = float(size)
    faces = [((0, 0, 0), (0, 0, s), (s, 0, 0)), ((0, s, 0), (s, 0, 0), (0, 0, s)),
             ((0, 0, 0), (0, s, 0), (0, 0, s)), ((s, 0, 0), (0, 0, s), (0, s, 0)),
             ((0, 0, 0), (s, 0, 0), (0, s, 0)), ((0, 0, s), (0, s, 0), (s, 0, 0))]
    centre = np.array([s / 2] * 3)
    for v0, e1, e2 in faces:
        # Tripar normal is cross(e2, e1) (surface/tripar.cc:52); make it
        # point into the box so the emitting front face is inside.
        n = np.cross(np.array(e2, float), np.array(e1, float))
        mid = np.array(v0, float) + 0.5 * (np.array(e1, float) + np.array(e2, float))
        if np.dot(n, centre - mid) < 0:
            e1, e2 = e2, e1
        sd.tripar(v0, e1, e2, g, parallelogram=True)
    cam = camera((s / 2, s / 2, s / 2), (0, 0, 1), (0, 1, 0), (1, 0, 0), 0.5, 0.5)
    return sd, cam


# ---------------------------------------------------------------------------
# Icosphere soup
# ---------------------------------------------------------------------------

def icosphere(level):
    """Unit icosphere: 20 * 4^level triangles, 10 * 4^level + 2 vertices."""
    t = (1.0 + 5 ** 0.5) / 2.0
    v = np.array([(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0),
                  (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
                  (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)], np.float64)
    f = np.array([(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11),
                  (1, 5, 9), (5, 11, 4), (11, 10, 2), (10, 7, 6), (7, 1, 8),
                  (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9),
                  (4, 9, 5), (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)],
                 np.int64)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    for _ in range(level):
        e = np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]])
        e.sort(axis=1)
        key = e[:, 0] * (len(v) + 1) + e[:, 1]
        uk, inv = np.unique(key, return_inverse=True)
        a, b = uk // (len(v) + 1), uk % (len(v) + 1)
        mid = v[a] + v[b]
        mid /= np.linalg.norm(mid, axis=1, keepdims=True)
        base = len(v)
        v = np.concatenate([v, mid])
        nf = len(f)
        m01 = base + inv[:nf]
        m12 = base + inv[nf:2 * nf]
        m20 = base + inv[2 * nf:]
        f = np.concatenate([
            np.stack([f[:, 0], m01, m20], 1), np.stack([f[:, 1], m12, m01], 1),
            np.stack([f[:, 2], m20, m12], 1), np.stack([m01, m12, m20], 1)])
    return v, f


def _value_noise(p, rng_seed, octaves=3):
    """Seeded 3-D value noise in [-1, 1]: hashed lattice values with
    smoothstep trilinear interpolation, `octaves` octaves at doubling
    frequency and halving amplitude (normalised)."""
    def lattice(ix, iy, iz, octave):
        h = (ix.astype(np.uint64) * np.uint64(0x8DA6B343) ^
             iy.astype(np.uint64) * np.uint64(0xD8163841) ^
             iz.astype(np.uint64) * np.uint64(0xCB1AB31F) ^
             np.uint64((rng_seed * 0x9E3779B1 + octave * 0x85EBCA77) & 0xFFFFFFFFFFFFFFFF))
        h ^= h >> np.uint64(13)
        h *= np.uint64(0x5BD1E995)
        h &= np.uint64(0xFFFFFFFF)
        h ^= h >> np.uint64(15)
        return (h & np.uint64(0xFFFFFF)).astype(np.float64) / float(0xFFFFFF) * 2.0 - 1.0

    total = np.zeros(len(p))
    amp, freq, norm = 1.0, 2.0, 0.0
    for o in range(octaves):
        q = p * freq + 17.0
        i = np.floor(q).astype(np.int64)
        fr = q - i
        s = fr * fr * (3 - 2 * fr)
        acc = np.zeros(len(p))
        for dx in (0, 1):
            wx = s[:, 0] if dx else 1 - s[:, 0]
            for dy in (0, 1):
                wy = s[:, 1] if dy else 1 - s[:, 1]
                for dz in (0, 1):
                    wz = s[:, 2] if dz else 1 - s[:, 2]
                    acc += wx * wy * wz * lattice(i[:, 0] + dx, i[:, 1] + dy,
                                                  i[:, 2] + dz, o)
        total += amp * acc
        norm += amp
        amp *= 0.5
        freq *= 2.0
    return total / norm


def gaussian_lobe_envmap(width, height, seed, n_lobes=6, sky=0.1,
                         max_peak=1e3):
    """Lat-long HDR map (row 0 = north): sky + sum_k A_k exp(k_k (mu_k.d - 1))."""
    rng = np.random.default_rng(seed)
    amps = np.exp(rng.uniform(math.log(20.0), math.log(max_peak), n_lobes))
    kappas = np.exp(rng.uniform(math.log(20.0), math.log(400.0), n_lobes))
    cols = rng.uniform(0.5, 1.0, (n_lobes, 3))
    z = rng.uniform(0.15, 1.0, n_lobes)          # upper hemisphere (map z)
    phi = rng.uniform(0, 2 * math.pi, n_lobes)
    r = np.sqrt(1 - z * z)
    mu = np.stack([r * np.cos(phi), r * np.sin(phi), z], 1)
    # texel centres -> map-space direction (geometry/spherical-coords.h)
    u = (np.arange(width) + 0.5) / width
    v = 1.0 - (np.arange(height) + 0.5) / height   # row 0 = top = north
    lng = (u - 0.5) * 2 * math.pi
    lat = (v - 0.5) * math.pi
    LNG, LAT = np.meshgrid(lng, lat)
    d = np.stack([np.cos(LNG) * np.cos(LAT), -np.sin(LNG) * np.cos(LAT),
                  np.sin(LAT)], -1)
    img = np.full((height, width, 3), sky, np.float64)
    for k in range(n_lobes):
        g = amps[k] * np.exp(kappas[k] * (d @ mu[k] - 1.0))
        img += g[..., None] * cols[k]
    return img.astype(np.float32)


def icosphere_soup(n_spheres=16, level=6, seed=1234, env_size=(1024, 512),
                   cube=20.0, radius=1.0, glass_fraction=0.2,
                   aspect=16.0 / 9.0):
    """The 16-icosphere triangle soup; returns (SceneDesc, camera)."""
    rng = np.random.default_rng(seed)
    sd = SceneDesc()
    ground = sd.lambert((0.5, 0.5, 0.5))
    base_v, base_f = icosphere(level)
    n_glass = int(round(glass_fraction * n_spheres))
    glass_ids = set(rng.choice(n_spheres, n_glass, replace=False).tolist())
    half = cube / 2
    g = 40.0
    sd.mesh(np.array([(-g, 0, -g), (g, 0, -g), (g, 0, g), (-g, 0, g)], np.float32),
            np.array([[0, 2, 1], [0, 3, 2]], np.uint32), ground)
    for i in range(n_spheres):
        c = np.array([rng.uniform(-half, half), rng.uniform(radius * 1.1, cube),
                      rng.uniform(-half, half)])
        if i in glass_ids:
            mat = sd.glass(1.5)
        else:
            mat = sd.cook_torrance(tuple(rng.uniform(0.1, 0.7, 3)),
                                   (0.5, 0.5, 0.5), float(rng.uniform(0.05, 0.5)),
                                   float(rng.uniform(1.3, 1.8)))
        disp = 1.0 + 0.05 * _value_noise(base_v, seed * 131 + i)
        v = c + radius * base_v * disp[:, None]
        sd.mesh(v.astype(np.float32), base_f.astype(np.uint32), mat,
                axis=(0.0, 1.0, 0.0))
    sd.envmap(gaussian_lobe_envmap(env_size[0], env_size[1], seed + 7))
    cam = look_at((0.0, 12.0, -34.0), (0.0, 9.0, 0.0), vfov_deg=45.0,
                  aspect=aspect)
    return sd, cam
