"""The direct-lighting integrator and the Phong material on the device
against the oracle (same counter-based stream, per path): direct
lighting at every hit with light_samples samples per light, recursion
into specular reflection and transmission with the roulette beyond
depth 5, sphere lights' extra draws from the sequential stream."""
import numpy as np
import pytest

import oracle
from snogray_amd import abi, native, scenes
from tests.test_gpu_render import check_paths

pytestmark = pytest.mark.gpu
DIRECT = abi.SGR_INTEG_DIRECT


def materials_scene():
    """Lambert, Phong and Cook-Torrance surfaces, a mirror with an
    underlying Lambert, two nested glass spheres, a sphere light, an area
    light and an envmap."""
    sd = scenes.SceneDesc()
    wall = sd.lambert((0.6, 0.6, 0.6))
    ph = sd.phong((0.3, 0.25, 0.2), (0.5, 0.5, 0.5), 25.0)
    ct = sd.cook_torrance((0.3, 0.2, 0.1), (0.6, 0.6, 0.6), 0.2, 1.6)
    mir = sd.mirror((0.9, 0.9, 0.9), ior=1.5, color=(0.2, 0.1, 0.1))
    sd.tripar((-5, 0, -5), (0, 0, 10), (10, 0, 0), ph, parallelogram=True)
    sd.tripar((-5, 0, 5), (10, 0, 0), (0, 6, 0), ct, parallelogram=True)
    sd.tripar((5, 0, -5), (0, 6, 0), (0, 0, 10), wall, parallelogram=True)
    sd.sphere((-1.5, 1, 1), 1.0, mir)
    sd.sphere((1.5, 1, 0.5), 1.0, sd.glass(1.5))
    sd.sphere((1.5, 1, 0.5), 0.5, sd.glass(1.8))
    sd.sphere((0, 4.5, 0), 0.5, sd.glow((20.0, 18.0, 15.0)))
    sd.tripar((-1, 5.5, -1), (2, 0, 0), (0, 0, 2), sd.glow((6.0, 6.0, 5.0)), parallelogram=True)
    sd.envmap(scenes.gaussian_lobe_envmap(64, 32, 5))
    return sd, scenes.look_at((0, 2, -8), (0, 1.5, 0), vfov_deg=45.0)


def _compare(ctx, sd, cam, p, name):
    sc, osc = native.Scene(ctx, sd), oracle.OracleScene(sd)
    got, gst = sc.render_samples(cam, p)
    _, ref, ost = osc.render(cam, p, samples=True)
    check_paths(got, ref, 1e-4, 0.999, name)
    for k in ("scene_intersect_calls", "shadow_tests", "illum_calls"):
        assert abs(gst[k] - ost[k]) <= 0.002 * ost[k], (k, gst[k], ost[k])
    assert gst["direct_overflows"] == ost["direct_overflows"] == 0
    return sc, osc


def test_direct_integrator_paths_match_oracle(ctx):
    sd, cam = materials_scene()
    p = abi.make_params(40, 32, 4, integrator=DIRECT, light_samples=3, seed=7)
    sc, osc = _compare(ctx, sd, cam, p, "direct integrator")
    # accumulated image (splat over the chunks) against the oracle's
    acc, _ = sc.render(cam, p)
    oacc, _, _ = osc.render(cam, p)
    fin = native.finalize(acc)[..., :3]
    ofin = native.finalize(oacc.astype(np.float32))[..., :3]
    assert np.mean(np.abs(fin - ofin) <= 1e-3 * (np.abs(ofin) + 1e-2)) > 0.99
    sc.close()


def test_direct_integrator_default_light_samples_on_cornell(ctx):
    """16 light samples (the default) on the Cornell box, and chunking:
    a render of several chunks equals one of a single chunk up to the
    order in which the splat sums a pixel's samples."""
    sd, cam = scenes.cornell_box()
    p = abi.make_params(32, 32, 4, integrator=DIRECT, seed=3)
    sc, _ = _compare(ctx, sd, cam, p, "direct integrator, Cornell")
    a, _ = sc.render(cam, p)
    q = abi.make_params(32, 32, 4, integrator=DIRECT, seed=3, paths_in_flight=36 * 36)
    b, _ = sc.render(cam, q)
    np.testing.assert_allclose(a, b, rtol=1e-5, atol=1e-6 * np.abs(a).max())
    sc.close()


def test_direct_integrator_on_instances(ctx):
    from tests.test_gpu_trace import instanced_scene
    sd, cam = instanced_scene()
    p = abi.make_params(32, 24, 2, integrator=DIRECT, light_samples=2)
    _compare(ctx, sd, cam, p, "direct integrator, instances")[0].close()


def test_phong_paths_match_oracle(ctx):
    """The path integrator over the same materials (Phong included)."""
    sd, cam = materials_scene()
    p = abi.make_params(40, 32, 8, max_path_len=8)
    _compare(ctx, sd, cam, p, "path integrator, Phong")[0].close()


def test_engine_and_guide_refuse_the_direct_integrator(ctx):
    sd, cam = scenes.cornell_box()
    sc = native.Scene(ctx, sd)
    p = abi.make_params(16, 16, 1, integrator=DIRECT)
    eng = native.Engine(ctx, path_capacity=1 << 20)
    with pytest.raises(native.SgrError, match="path integrator only"):
        eng.submit(sc, cam, p)
    g = native.Guide(ctx)
    with pytest.raises(native.SgrError, match="path integrator only"):
        sc.render_guided(cam, p, g)
    g.close()
    p.integrator = 7
    with pytest.raises(native.SgrError, match="unknown integrator"):
        sc.render(cam, p)
    sc.close()
