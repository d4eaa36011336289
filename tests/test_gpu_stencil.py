"""Stencils on the device against the oracle (material/stencil.cc): the
opacity-weighted underlying BSDF and the straight-through branch, shadow
rays through them (transmittance max (1 - opacity, 0), once per
surface), fully opaque and nearly transparent stencils, and stencils
over glass, a mirror, a glow and Cook-Torrance, under both integrators,
in the ray-query API and in the job engine."""
import numpy as np
import pytest

import oracle
from snogray_amd import abi, batch, native, scenes
from tests import helpers
from tests.test_gpu_render import check_paths

pytestmark = pytest.mark.gpu


def stencil_scene():
    """A small stencil pane (listed in several octree nodes) between a
    floor and an area light, stencil balls over Cook-Torrance, glass and
    a glow-wrapped Lambert, a fully opaque stencil over a mirror, a
    nearly transparent one over Lambert, a thin glass pane, a sphere
    light and an envmap."""
    sd = scenes.SceneDesc()
    floor = sd.lambert((0.6, 0.6, 0.6))
    sd.tripar((-50, 0, -50), (100, 0, 0), (0, 0, 100), floor, parallelogram=True)
    sd.tripar((-5, 2.2, -5), (10, 0, 0), (0, 0, 10),
              sd.stencil((0.3, 0.5, 0.7), sd.lambert((0.8, 0.4, 0.2))), parallelogram=True)
    sd.tripar((-0.6, 3.5, -0.4), (0, 0, 0.9), (1.1, 0, 0), sd.glow((8.0, 8.0, 8.0)),
              parallelogram=True)
    sd.sphere((-1.2, 0.8, 0.5), 0.8,
              sd.stencil((0.6, 0.6, 0.6), sd.cook_torrance((0.5, 0.3, 0.2), (0.4, 0.4, 0.4), 0.2, 1.6)))
    sd.sphere((1.3, 0.9, 0.2), 0.9, sd.stencil((0.5, 0.4, 0.3), sd.glass(1.5)))
    sd.sphere((0.0, 0.5, -1.4), 0.45, sd.stencil((1.0, 1.0, 1.0), sd.mirror((0.9, 0.9, 0.9))))
    sd.sphere((-2.4, 0.4, -0.8), 0.4,
              sd.stencil((0.7, 0.2, 0.5), sd.glow((0.5, 0.5, 0.5), sd.lambert((0.3, 0.6, 0.3)))))
    sd.tripar((0.7, 0.2, -0.3), (1.2, 0, 0), (0, 1.4, 0),
              sd.stencil((0.0005, 0.0005, 0.0005), sd.lambert((0.9, 0.9, 0.9))), parallelogram=True)
    sd.tripar((-2.0, 0.1, 1.5), (1.0, 0, 0), (0, 1.2, 0), sd.thin_glass((0.9, 0.8, 0.9), 1.5),
              parallelogram=True)
    sd.tripar((-5, 0, 4), (10, 0, 0), (0, 5, 0), sd.lambert((0.4, 0.5, 0.6)), parallelogram=True)
    sd.sphere((2.5, 4.0, -1.0), 0.3, sd.glow((20.0, 18.0, 15.0)))
    sd.envmap(scenes.gaussian_lobe_envmap(64, 32, 7))
    return sd, scenes.look_at((0, 1.6, -6), (0, 1.0, 0), vfov_deg=50.0)


def _compare(ctx, sd, cam, p, name, min_frac=0.999):
    sc, osc = native.Scene(ctx, sd), oracle.OracleScene(sd)
    got, gst = sc.render_samples(cam, p)
    _, ref, ost = osc.render(cam, p, samples=True)
    check_paths(got, ref, 1e-4, min_frac, name)
    for k in ("scene_intersect_calls", "shadow_tests", "illum_calls", "surf_tests_shadow"):
        assert abs(gst[k] - ost[k]) <= 0.002 * ost[k], (k, gst[k], ost[k])
    assert gst["partial_overflows"] == 0
    return sc, osc


def test_stencil_paths_match_oracle(ctx):
    sd, cam = stencil_scene()
    p = abi.make_params(40, 32, 8, max_path_len=8, direct_samples=2)
    sc, osc = _compare(ctx, sd, cam, p, "stencil, path integrator")
    q = abi.make_params(40, 32, 8, max_path_len=8, direct_samples=2,
                        flags=abi.SGR_RENDER_WAVEFRONT_ONLY)
    a, _ = sc.render(cam, p)
    b, _ = sc.render(cam, q)
    np.testing.assert_array_equal(a, b)   # k_finish and the per-bounce kernels agree
    sc.close()


def test_stencil_direct_integrator_matches_oracle(ctx):
    """One sequential stream feeds every roulette draw of a sample in the
    direct integrator, so a branch taken on one side only (a last-ulp
    difference at a total-internal-reflection boundary inside the glass
    ball) shifts the sample's later draws: 0.998, as for thin glass."""
    sd, cam = stencil_scene()
    p = abi.make_params(40, 32, 2, integrator=abi.SGR_INTEG_DIRECT, light_samples=3, seed=4)
    _compare(ctx, sd, cam, p, "stencil, direct integrator", 0.998)[0].close()


def test_occluded_queries_through_stencils_bit_exact(ctx):
    sd, _ = stencil_scene()
    sc, osc = native.Scene(ctx, sd), oracle.OracleScene(sd)
    lo, hi, _ = sc.bounds()
    b = helpers.random_rays(np.maximum(lo, -6), np.minimum(hi, 6), 60000, seed=11)
    g, gst = sc.trace_occluded(b)
    o, (onodes, otests) = osc.trace_occluded(b)
    assert np.array_equal(g, o)
    assert gst["surf_tests_shadow"] == otests
    c, _ = osc.trace_closest(b)
    hit = c.prim != abi.MISS
    assert (hit & (g == 0)).sum() > 500   # rays through stencils (or thin glass) only
    sc.close()


def test_stencil_job_in_a_batch_equals_alone(ctx):
    sd, cam = stencil_scene()
    sd2, cam2 = scenes.cornell_box()
    s1, s2 = native.Scene(ctx, sd), native.Scene(ctx, sd2)
    jobs = [(s1, cam, abi.make_params(32, 24, 4, max_path_len=6, seed=3)),
            (s2, cam2, abi.make_params(24, 24, 4, max_path_len=6, seed=5))]
    alone = [sc.render(c, p)[0] for sc, c, p in jobs]
    rep = batch.run_batch(ctx, jobs, path_capacity=1 << 20, keep_results=True)
    for i, a in enumerate(alone):
        np.testing.assert_array_equal(rep["results"][i], a)
    s1.close()
    s2.close()


def test_stencil_nesting_refused(ctx):
    sd = scenes.SceneDesc()
    s1 = sd.stencil((0.5, 0.5, 0.5), sd.lambert((0.5, 0.5, 0.5)))
    sd.sphere((0, 0, 0), 1.0, sd.stencil((0.5, 0.5, 0.5), sd.glow((1, 1, 1), s1)))
    with pytest.raises(RuntimeError, match="stencil"):
        native.Scene(ctx, sd)
    with pytest.raises(RuntimeError, match="stencil"):
        oracle.OracleScene(sd)
