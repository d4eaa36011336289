"""Vertex normals on meshes (surface/mesh.cc:620-641): the shading normal
is the barycentric interpolation (1 - u - v, u, v) of the three vertex
normals, normalised, and the geometric normal is flipped to its side.
Pinned on the oracle by a closed form for the interpolation, by the
flip making a mesh of either handedness render the same, and by face
normals reproducing the mesh without normals."""
import math

import numpy as np
import pytest

import oracle
from snogray_amd import abi, scenes


def unit(v):
    v = np.asarray(v, float)
    return v / np.linalg.norm(v, axis=-1, keepdims=True)


FLOOR_V = np.array([(-3, 0, -3), (3, 0, -3), (3, 0, 3), (-3, 0, 3)], np.float32)
FLOOR_T = np.array([[0, 2, 1], [0, 3, 2]], np.uint32)
# tilted vertex normals, each within 20 degrees of +y
FLOOR_N = unit([(0.3, 1, 0.1), (-0.2, 1, 0.25), (0.05, 1, -0.3), (-0.25, 1, -0.1)])


def floor_handedness():
    """left_handed flag that makes FLOOR_T's geometric normal +y."""
    a, b, c = FLOOR_V[FLOOR_T[0]].astype(float)
    return bool(np.cross(b - a, c - a)[1] < 0)  # raw = left ? e2 x e1 : e1 x e2


def barycentric(pts, tri):
    """(u, v) of points in the plane of TRI (corner tri[0], edges to
    tri[1] and tri[2])."""
    a, b, c = tri.astype(float)
    e1, e2 = b - a, c - a
    d = pts - a
    m = np.array([[e1 @ e1, e1 @ e2], [e1 @ e2, e2 @ e2]])
    rhs = np.stack([d @ e1, d @ e2], 1)
    uv = np.linalg.solve(m, rhs.T).T
    return uv[:, 0], uv[:, 1]


def test_shading_normal_is_the_barycentric_interpolation():
    """Lambert floor with tilted vertex normals under a small sphere
    light, direct lighting only: each sample's radiance is
    rho Le (r/D)^2 (n_s . w), with n_s = unit((1-u-v) n0 + u n1 + v n2)
    at the hit's barycentrics -- the cosine of the BSDF is taken about
    the interpolated normal, not the face normal."""
    Le, rad, c = 10.0, 0.4, np.array([0.0, 3.0, 0.0])
    sd = scenes.SceneDesc()
    floor = sd.lambert((0.5, 0.5, 0.5))
    sd.mesh(FLOOR_V, FLOOR_T, floor, left_handed=floor_handedness(), normals=FLOOR_N)
    sd.sphere(tuple(c), rad, sd.glow((Le, Le, Le)))
    cam = scenes.camera((0.0, 1.5, 0.0), (0, -1, 0), (0, 0, 1), (1, 0, 0), 0.9, 0.9)
    W = H = 24
    p = abi.make_params(W, H, 64, max_path_len=0, min_path_len=0, seed=5)
    _, smp, _ = oracle.OracleScene(sd).render(cam, p, samples=True)
    sx, sy = smp[:, 0].astype(float), smp[:, 1].astype(float)
    f, r, up = (np.asarray(x, float) for x in (cam.forward, cam.right, cam.up))
    d = f + np.outer(2 * sx / W - 1, r) * cam.tan_half_fov_x + \
        np.outer(2 * (H - sy) / H - 1, up) * cam.tan_half_fov_y
    pts = np.asarray(cam.pos, float) + d * (-1.5 / d[:, 1])[:, None]
    ns = np.zeros_like(pts)
    for tri in FLOOR_T:
        u, v = barycentric(pts, FLOOR_V[tri])
        inside = (u >= -1e-6) & (v >= -1e-6) & (u + v <= 1 + 1e-6)
        n0, n1, n2 = FLOOR_N[tri]
        ns[inside] = unit((1 - u - v)[inside, None] * n0 + u[inside, None] * n1
                          + v[inside, None] * n2)
    dv = c - pts
    D = np.linalg.norm(dv, axis=1)
    expect = 0.5 * Le * (rad / D) ** 2 * np.maximum((ns * dv).sum(1) / D, 0)
    flat = 0.5 * Le * (rad / D) ** 2 * (dv[:, 1] / D)
    got = smp[:, 2].astype(float)
    diff = got - expect
    se = diff.std() / math.sqrt(len(diff))
    assert abs(diff.mean()) < 3 * se + 1e-4 * expect.mean()
    # the face normal's cosine would be off by many standard errors here,
    # and so would the vertex weights in the wrong order
    for region in (pts[:, 0] < -1, pts[:, 0] > 1, pts[:, 2] < -1, pts[:, 2] > 1):
        dd = diff[region]
        assert abs(dd.mean()) < 4 * dd.std() / math.sqrt(len(dd)) + 1e-4 * expect.mean()
        df = (got - flat)[region]
        assert abs(df.mean()) > 8 * df.std() / math.sqrt(len(df))


def _glass_ball(left_handed, normals, level=3):
    v, t = scenes.icosphere(level)
    v = np.asarray(v, np.float32)
    sd = scenes.SceneDesc()
    sd.mesh(v, t, sd.glass(1.5), left_handed=left_handed,
            normals=(v if normals else None))
    sd.envmap(scenes.gaussian_lobe_envmap(64, 32, 3))
    return sd


def _outward_handedness():
    v, t = scenes.icosphere(3)
    a, b, c = np.asarray(v, float)[t[0]]
    raw = np.cross(b - a, c - a)          # right-handed raw normal
    return bool(raw @ (a + b + c) < 0)   # left-handed if that points inward


def test_geometric_normal_follows_the_vertex_normals():
    """A glass ball whose faces are wound the wrong way renders like the
    correctly wound one once its vertex normals point outward: the
    geometric normal is flipped to the shading normal's side, so entering
    and leaving the glass (the back-face flag) come out the same, bit for
    bit; without the normals the wrong winding swaps them."""
    cam = scenes.look_at((0, 0.2, -4), (0, 0, 0), vfov_deg=35.0)
    p = abi.make_params(24, 24, 4, max_path_len=8, seed=9)
    good = _outward_handedness()
    a = oracle.OracleScene(_glass_ball(good, True)).render(cam, p, samples=True)[1]
    b = oracle.OracleScene(_glass_ball(not good, True)).render(cam, p, samples=True)[1]
    assert np.array_equal(a, b)
    c = oracle.OracleScene(_glass_ball(not good, False)).render(cam, p, samples=True)[1]
    d = oracle.OracleScene(_glass_ball(good, False)).render(cam, p, samples=True)[1]
    assert not np.allclose(c[:, 2:5], d[:, 2:5], rtol=1e-3, atol=1e-3)


def _unshared(sd0, normals):
    """SD0 with every mesh triangle on its own three vertices, and, with
    NORMALS, each vertex normal the triangle's own face normal."""
    sd = scenes.SceneDesc()
    sd.materials = [dict(m) for m in sd0.materials]
    for m in sd0.meshes:
        v = m["v"][m["t"].ravel()]
        t = np.arange(len(v), dtype=np.uint32).reshape(-1, 3)
        a, b, c = v[0::3].astype(float), v[1::3].astype(float), v[2::3].astype(float)
        raw = np.cross(c - a, b - a) if m["left_handed"] else np.cross(b - a, c - a)
        n = np.repeat(unit(raw), 3, axis=0) if normals else None
        sd.mesh(v, t, m["material"], tri_materials=m["tm"], axis=m["axis"],
                left_handed=m["left_handed"], normals=n)
    for tp in sd0.tripars:
        sd.tripar(tp["v0"], tp["e1"], tp["e2"], tp["material"], parallelogram=tp["parallelogram"])
    return sd


def test_face_normals_reproduce_the_mesh_without_normals():
    """Vertex normals equal to each face's own normal: the interpolated
    normal is the face normal up to rounding, so every path agrees with
    the render without normals to 1e-5."""
    sd0, cam = scenes.cornell_box()
    p = abi.make_params(32, 32, 8, max_path_len=8)
    ref = oracle.OracleScene(_unshared(sd0, False)).render(cam, p, samples=True)[1]
    got = oracle.OracleScene(_unshared(sd0, True)).render(cam, p, samples=True)[1]
    rel = np.abs(got[:, 2:] - ref[:, 2:]) / (np.abs(ref[:, 2:]) + 1e-3)
    assert np.mean(rel.max(1) <= 1e-5) > 0.999
