"""Two-level instancing on the oracle (surface/instance.cc:87-146,
surface/model.cc:26-47): a model's own octree searched with the ray taken
to model space, the hit frame mapped back (the normal by the inverse
transpose).  Pinned by an identity instance reproducing the geometry
placed directly, a rigid-plus-scale instance matching the transformed
geometry, and the analytic normal of a non-uniformly scaled sphere."""
import math

import numpy as np
import pytest

import oracle
from snogray_amd import abi, native, scenes
from tests import helpers

I34 = np.hstack([np.eye(3), np.zeros((3, 1))])


def box_model_scene(instanced, xform=I34):
    """Cornell box whose two boxes are one model placed by XFORM
    (instanced) or the same triangles given directly (transformed on the
    host in float64)."""
    sd0, cam = scenes.cornell_box()
    sd = scenes.SceneDesc()
    sd.materials = [dict(m) for m in sd0.materials]
    walls = sd0.meshes[0]
    sd.mesh(walls["v"], walls["t"], walls["material"], tri_materials=walls["tm"])
    for tp in sd0.tripars:
        sd.tripar(tp["v0"], tp["e1"], tp["e2"], tp["material"], parallelogram=tp["parallelogram"])
    boxes = sd0.meshes[1:]
    if instanced:
        mi, m = sd.model()
        for b in boxes:
            m.mesh(b["v"], b["t"], b["material"])
        sd.instance(mi, xform)
    else:
        A, t = np.asarray(xform)[:, :3], np.asarray(xform)[:, 3]
        for b in boxes:
            v = (b["v"].astype(np.float64) @ A.T + t).astype(np.float32)
            sd.mesh(v, b["t"], b["material"])
    return sd, cam


def test_identity_instance_equals_direct_geometry():
    """Identity transform: the model-space ray is the ray itself, so every
    hit has the same t, u, v and surface as with the boxes in the scene
    (model triangle k = scene triangle 10 + k)."""
    si, cam = box_model_scene(True)
    sd, _ = box_model_scene(False)
    batch = helpers.random_rays((0.2, 0.2, 0.2), (5.3, 5.3, 5.3), 20000, seed=3)
    hi, _ = oracle.OracleScene(si).trace_closest(batch)
    hd, _ = oracle.OracleScene(sd).trace_closest(batch)
    box_hit = (hd.prim != abi.MISS) & (hd.prim >= 10) & (hd.prim < 34)
    inst_hit = hi.inst == 0
    assert box_hit.sum() > 2000
    # same nearest distance everywhere; the boxes' bottoms lie in the floor's
    # plane, so a few rays meet two surfaces at the same t and either may
    # win (the search order differs), nothing else differs
    assert np.array_equal(hi.t, hd.t)
    same = np.where(inst_hit, hi.prim + 10, hi.prim)
    same = np.where(hi.prim == abi.MISS, abi.MISS, same)
    same = np.where(~inst_hit & (hi.prim == 10), 34, same)  # the light, after the boxes
    agree = same == hd.prim
    assert agree.mean() > 0.995
    assert np.array_equal(hi.u[agree], hd.u[agree]) and np.array_equal(hi.v[agree], hd.v[agree])
    assert np.all(inst_hit[box_hit & agree])


def rot(axis, ang):
    axis = np.asarray(axis, float) / np.linalg.norm(axis)
    K = np.array([[0, -axis[2], axis[1]], [axis[2], 0, -axis[0]], [-axis[1], axis[0], 0]])
    return np.eye(3) + math.sin(ang) * K + (1 - math.cos(ang)) * K @ K


def test_rotated_scaled_instance_matches_transformed_geometry():
    """Rotation, uniform scale and translation: hits on the instanced
    model match the transformed triangles given directly (same surfaces,
    t within float rounding of the two transforms)."""
    A = 0.8 * rot((0.2, 1.0, 0.1), 0.6)
    X = np.hstack([A, np.array([[0.3], [0.1], [-0.2]])])
    si, _ = box_model_scene(True, X)
    sd, _ = box_model_scene(False, X)
    batch = helpers.random_rays((0.2, 0.2, 0.2), (5.3, 5.3, 5.3), 20000, seed=4)
    hi, _ = oracle.OracleScene(si).trace_closest(batch)
    hd, _ = oracle.OracleScene(sd).trace_closest(batch)
    box_hit = (hd.prim != abi.MISS) & (hd.prim >= 10) & (hd.prim < 34)
    assert box_hit.sum() > 1500
    same = (hi.prim[box_hit] == hd.prim[box_hit] - 10) & (hi.inst[box_hit] == 0)
    assert same.mean() > 0.995
    rel = np.abs(hi.t[box_hit] - hd.t[box_hit]) / hd.t[box_hit]
    assert np.median(rel) < 1e-6 and rel.max() < 1e-3


def test_scaled_sphere_normal_is_the_inverse_transpose():
    """A unit sphere model scaled to an ellipsoid (2, 1, 0.5), rotated and
    moved: hits lie on the ellipsoid and the shading normal is the
    analytic gradient (x/a^2, y/b^2, z/c^2) taken to world space -- the
    inverse transpose, not the transform itself."""
    S = np.diag([2.0, 1.0, 0.5])
    Rm = rot((1.0, 0.3, -0.2), 0.9)
    A, t = Rm @ S, np.array([1.0, -0.5, 2.0])
    sd = scenes.SceneDesc()
    mi, m = sd.model()
    m.sphere((0, 0, 0), 1.0, sd.lambert((0.5, 0.5, 0.5)))
    sd.instance(mi, np.hstack([A, t[:, None]]))
    osc = oracle.OracleScene(sd)
    rng = np.random.default_rng(2)
    o = t + rng.normal(size=(4000, 3)) * 0.3 + np.array([0, 0, -8.0])
    d = (t + rng.uniform(-1.2, 1.2, (4000, 3)) * [2, 1, 1]) - o
    h = osc.isec_probe(o, d)
    hit = h["prim"] >= 0
    assert hit.sum() > 1000 and np.all(h["inst"][hit] == 0)
    p = h["origin"][hit]
    q = np.linalg.solve(A, (p - t).T).T          # model space: on the unit sphere
    assert np.allclose(np.linalg.norm(q, axis=1), 1, atol=2e-5)
    n_ref = np.linalg.inv(A).T @ q.T              # gradient of |A^-1 (p - t)|^2
    n_ref = (n_ref / np.linalg.norm(n_ref, axis=0)).T
    n = h["normal"][hit]
    assert np.allclose(np.abs((n * n_ref).sum(1)), 1, atol=1e-5)
    # the transform itself would tilt them visibly
    n_bad = (A @ q.T).T
    n_bad /= np.linalg.norm(n_bad, axis=1, keepdims=True)
    assert np.mean(np.abs((n * n_bad).sum(1)) < 0.99) > 0.5


def test_instanced_render_matches_direct_geometry():
    """Identity instance: the shading frames differ from the direct ones
    by the renormalisation of already unit axes (an ulp), so the paths
    agree to 1e-4."""
    si, cam = box_model_scene(True)
    sd, _ = box_model_scene(False)
    p = abi.make_params(24, 24, 4, max_path_len=8)
    a = oracle.OracleScene(si).render(cam, p, samples=True)[1]
    b = oracle.OracleScene(sd).render(cam, p, samples=True)[1]
    rel = np.abs(a[:, 2:] - b[:, 2:]) / (np.abs(b[:, 2:]) + 1e-3)
    assert np.mean(rel.max(1) <= 1e-4) > 0.99


def test_host_octree_with_instances_matches_oracle():
    """The library's host build places each instance by its transformed
    model box exactly as the oracle does (same scene octree)."""
    A = 0.8 * rot((0.2, 1.0, 0.1), 0.6)
    si, _ = box_model_scene(True, np.hstack([A, np.array([[0.3], [0.1], [-0.2]])]))
    mi, m = si.model()
    m.sphere((0, 0, 0), 1.0, 0)
    for k in range(5):
        si.instance(mi, np.hstack([np.diag([0.2, 0.3, 0.25]), np.array([[1.0 + k], [2.0], [3.0]])]))
    n1, s1, o1, _ = native.octree_build(si)
    n2, s2, o2 = oracle.OracleScene(si).octree()
    assert np.array_equal(n1, n2) and np.array_equal(s1, s2)
    assert np.array_equal(o1.astype(np.float64), o2)
    # surface ids: the direct primitives first, then the 6 instances
    n_prims = len(si.meshes[0]["t"]) + len(si.tripars)
    assert set(range(n_prims, n_prims + 6)) <= set(s1.tolist())
