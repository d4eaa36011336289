"""Ray-query parity on the GPU: closest hits (primitive id, t, u, v) and
any-hit occlusion must equal the CPU oracle bit for bit, and the octree
walk must visit the same nodes and test the same surfaces."""
import numpy as np
import pytest

import oracle
from snogray_amd import abi, native, scenes
from tests import helpers

pytestmark = pytest.mark.gpu


def scene_pair(ctx, name):
    if name == "cornell":
        sd, cam = scenes.cornell_box()
    elif name == "furnace":
        sd, cam = scenes.furnace_box()
    elif name == "icosoup_small":
        sd, cam = helpers.small_soup(level=5, n_spheres=3)
    elif name == "tri_soup":
        sd, cam = helpers.triangle_soup(20000, seed=4, extent=15.0, size=0.4), None
    elif name == "spheres":
        sd = scenes.SceneDesc()
        m = sd.lambert((0.5, 0.5, 0.5))
        rng = np.random.default_rng(5)
        for c in rng.uniform(-5, 5, (40, 3)):
            sd.sphere(tuple(c), float(rng.uniform(0.2, 1.0)), m)
        sd.tripar((-6, -6, -6), (12, 0, 0), (0, 0, 12), m, parallelogram=True)
        cam = None
    else:
        raise KeyError(name)
    return sd, cam, native.Scene(ctx, sd), oracle.OracleScene(sd)


def assert_hits_equal(g, o):
    assert np.array_equal(g.prim, o.prim)
    hit = o.prim != abi.MISS
    assert np.array_equal(g.t[hit].view(np.uint32), o.t[hit].view(np.uint32))
    assert np.array_equal(g.u[hit].view(np.uint32), o.u[hit].view(np.uint32))
    assert np.array_equal(g.v[hit].view(np.uint32), o.v[hit].view(np.uint32))


@pytest.mark.parametrize("name", ["cornell", "furnace", "icosoup_small", "tri_soup", "spheres"])
def test_closest_and_occluded_bit_exact(ctx, name):
    sd, cam, sc, osc = scene_pair(ctx, name)
    lo, hi, _ = sc.bounds()
    pad = 0.2 * (hi - lo)
    batches = [helpers.random_rays(lo - pad, hi + pad, 100000, seed=1)]
    if cam is not None:
        batches.append(helpers.camera_rays(cam, 128, 96, seed=2))
    for b in batches:
        g, gst = sc.trace_closest(b)
        o, (onodes, otests) = osc.trace_closest(b)
        assert_hits_equal(g, o)
        # same walk: the surfaces tested are exactly the recursive search's;
        # node calls agree up to siblings the device prunes before their
        # turn (counted when pruned, space/octree.cc:182)
        assert gst["surf_tests_closest"] == otests
        assert abs(gst["node_visits_closest"] - onodes) <= 0.1 * onodes
        for sb in (helpers.shadow_batch(b, o), helpers.shadow_batch(b, o, frac=1.001)):
            go, _ = sc.trace_occluded(sb)
            oo, _ = osc.trace_occluded(sb)
            assert np.array_equal(go, oo)


def test_full_soup_bit_exact(ctx):
    """The 1,310,722-triangle workload scene (octree depth 19)."""
    sd, cam = scenes.icosphere_soup()
    sc, osc = native.Scene(ctx, sd), oracle.OracleScene(sd)
    st = sc.octree_stats()
    assert st["max_depth"] == osc.octree_stats()["max_depth"]
    lo, hi, _ = sc.bounds()
    for b in (helpers.camera_rays(cam, 320, 180, seed=3),
              helpers.random_rays(lo, hi, 150000, seed=4)):
        g, gst = sc.trace_closest(b)
        o, (onodes, otests) = osc.trace_closest(b)
        assert_hits_equal(g, o)
        assert abs(gst["node_visits_closest"] - onodes) <= 0.1 * onodes
        assert gst["surf_tests_closest"] == otests
        go, _ = sc.trace_occluded(helpers.shadow_batch(b, o))
        oo, _ = osc.trace_occluded(helpers.shadow_batch(b, o))
        assert np.array_equal(go, oo)
    sc.close()


def test_exported_octree_matches_oracle(ctx):
    sd, _ = helpers.small_soup(level=4, n_spheres=3)
    sc = native.Scene(ctx, sd)
    n1, s1, o1 = sc.octree()
    n2, s2, o2 = oracle.OracleScene(sd).octree()
    assert np.array_equal(n1, n2) and np.array_equal(s1, s2)


def test_brute_force_accel_matches_octree(ctx):
    """accel=triv (TrivSpace) against the octree on the GPU: identical
    except near-ties in t."""
    sd = helpers.triangle_soup(800, seed=8, extent=5.0, size=0.6)
    sdt = helpers.triangle_soup(800, seed=8, extent=5.0, size=0.6, accel=abi.SGR_ACCEL_TRIV)
    a, b_ = native.Scene(ctx, sd), native.Scene(ctx, sdt)
    rays = helpers.random_rays((-6, -6, -6), (6, 6, 6), 20000, seed=9)
    ga, _ = a.trace_closest(rays)
    gb, _ = b_.trace_closest(rays)
    ob, _ = oracle.OracleScene(sdt).trace_closest(rays)
    assert_hits_equal(gb, ob)
    diff = ga.prim != gb.prim
    assert np.all(np.abs(ga.t[diff] - gb.t[diff]) <= 1e-6 * np.abs(gb.t[diff]) + 1e-7)


def test_edge_rays(ctx):
    """Axis-parallel directions (inverse 1e9 sentinel, space/octree.cc:119-135),
    rays starting outside the octree, empty and one-ray batches."""
    sd, cam = scenes.cornell_box()
    sc, osc = native.Scene(ctx, sd), oracle.OracleScene(sd)
    o = np.array([[2.75, 2.75, 2.75], [2.75, 2.75, -5], [-3, 1, 1], [2.0, 0.5, 2.0],
                  [2.75, 5.49, 2.75], [0, 0, 0]], np.float32)
    d = np.array([[1, 0, 0], [0, 0, 1], [1, 0, 0], [0, -1, 0], [0, 1, 0], [1, 1, 1]],
                 np.float32)
    b = abi.RayBatch(np.repeat(o, 6, 0), np.tile(d, (6, 1)), np.float32(1e-3), np.float32(100))
    assert_hits_equal(sc.trace_closest(b)[0], osc.trace_closest(b)[0])
    one = abi.RayBatch(o[:1], d[:1], np.float32(0), np.float32(1e30))
    assert_hits_equal(sc.trace_closest(one)[0], osc.trace_closest(one)[0])
    empty = abi.RayBatch(np.zeros((0, 3), np.float32), np.zeros((0, 3), np.float32),
                         np.float32(0), np.float32(1))
    h, _ = sc.trace_closest(empty)
    assert len(h.prim) == 0


def instanced_scene():
    """Cornell box whose two boxes are one model placed twice (rotated and
    scaled, then mirrored), with an ellipsoid made of a scaled unit-sphere
    model, and a model holding its own triangle soup placed 20 times."""
    from tests.test_oracle_instances import box_model_scene, rot
    A = 0.8 * rot((0.2, 1.0, 0.1), 0.6)
    sd, cam = box_model_scene(True, np.hstack([A, np.array([[0.3], [0.1], [-0.2]])]))
    sd.instance(0, np.hstack([np.diag([-0.5, 0.6, 0.5]), np.array([[4.5], [0.0], [3.0]])]))
    mi, m = sd.model()
    m.sphere((0, 0, 0), 1.0, 0)
    sd.instance(mi, np.hstack([rot((1, 0.3, -0.2), 0.9) @ np.diag([0.6, 0.3, 0.15]),
                               np.array([[2.0], [4.0], [2.5]])]))
    soup = helpers.triangle_soup(2000, seed=9, extent=1.0, size=0.15)
    mj, ms = sd.model()
    ms.mesh(soup.meshes[0]["v"], soup.meshes[0]["t"], 0)
    rng = np.random.default_rng(8)
    for k in range(20):
        X = rot(rng.normal(size=3), rng.uniform(0, 3)) * rng.uniform(0.2, 0.5)
        sd.instance(mj, np.hstack([X, rng.uniform(0.8, 4.8, (3, 1))]))
    return sd, cam


def test_instanced_scene_bit_exact(ctx):
    """Two-level search: closest hits (primitive within its model, the
    instance, t, u, v), the surfaces tested and any-hit occlusion equal
    the oracle's bit for bit."""
    sd, cam = instanced_scene()
    sc, osc = native.Scene(ctx, sd), oracle.OracleScene(sd)
    lo, hi, _ = sc.bounds()
    for b in (helpers.random_rays(lo, hi, 100000, seed=5), helpers.camera_rays(cam, 128, 96, seed=6)):
        g, gst = sc.trace_closest(b)
        o, (onodes, otests) = osc.trace_closest(b)
        assert_hits_equal(g, o)
        hit = o.prim != abi.MISS
        assert np.array_equal(g.inst[hit], o.inst[hit])
        assert (o.inst[hit] != abi.MISS).sum() > 0.1 * hit.sum()
        assert gst["surf_tests_closest"] == otests
        assert abs(gst["node_visits_closest"] - onodes) <= 0.1 * onodes
        for sb in (helpers.shadow_batch(b, o), helpers.shadow_batch(b, o, frac=1.001)):
            go, _ = sc.trace_occluded(sb)
            oo, _ = osc.trace_occluded(sb)
            assert np.array_equal(go, oo)
    sc.close()
