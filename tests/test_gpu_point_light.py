"""Point lights on the device against the oracle (light/point-light.cc):
per-path radiance of the path tracer and of the direct-lighting
integrator with an omni light and a spotlight with a fringe beside the
Cornell box's area light, the diffuse-sphere closed form on the device,
photon maps shot from point lights, and a job in a batch equal to the job
alone."""
import math

import numpy as np
import pytest

import oracle
from snogray_amd import abi, batch, native, scenes
from tests.test_gpu_photon import _compare_maps
from tests.test_gpu_render import check_paths

pytestmark = pytest.mark.gpu


def lit_cornell():
    sd, cam = scenes.cornell_box(light_emission=(4.0, 3.0, 1.0))
    sd.sphere((-0.1, 0.9, 0.2), 0.55, sd.glass(1.5))
    sd.point_light((1.2, 3.8, 1.5), (6.0, 5.0, 4.0))
    sd.point_light((-1.5, 4.5, -0.5), (9.0, 9.0, 12.0), angle=1.1, dir=(0.5, -1.0, 0.4),
                   fringe_angle=0.35)
    return sd, cam


def _compare(ctx, sd, cam, p, name):
    sc, osc = native.Scene(ctx, sd), oracle.OracleScene(sd)
    got, gst = sc.render_samples(cam, p)
    _, ref, ost = osc.render(cam, p, samples=True)
    check_paths(got, ref, 1e-4, 0.99, name)
    for k in ("scene_intersect_calls", "shadow_tests", "illum_calls"):
        assert abs(gst[k] - ost[k]) <= 0.002 * ost[k] + 2, (k, gst[k], ost[k])
    return sc


def test_point_light_paths_match_oracle(ctx):
    sd, cam = lit_cornell()
    p = abi.make_params(40, 32, 8, max_path_len=8, direct_samples=2)
    sc = _compare(ctx, sd, cam, p, "point lights, path tracer")
    # the persistent kernel finishing short queues and the per-bounce
    # kernels give the same image
    q = abi.make_params(40, 32, 8, max_path_len=8, direct_samples=2,
                        flags=abi.SGR_RENDER_WAVEFRONT_ONLY)
    a, _ = sc.render(cam, p)
    b, _ = sc.render(cam, q)
    np.testing.assert_array_equal(a, b)
    sc.close()


def test_point_light_direct_integrator_matches_oracle(ctx):
    sd, cam = lit_cornell()
    p = abi.make_params(40, 32, 4, integrator=abi.SGR_INTEG_DIRECT, light_samples=2)
    _compare(ctx, sd, cam, p, "point lights, direct integrator").close()


def test_point_light_in_a_diffuse_sphere(ctx):
    """Every camera sample is rho I / (pi R^2) (1 + rho + ... + rho^M)
    without Russian roulette (min path length = max = M)."""
    R, rho, M, I = 2.0, 0.5, 20, (3.0, 2.0, 1.0)
    sd = scenes.SceneDesc()
    sd.sphere((0.0, 0.0, 0.0), R, sd.lambert((rho,) * 3))
    sd.point_light((0.0, 0.0, 0.0), I)
    cam = scenes.camera((0.3, 0.1, -0.2), (0, 0, 1), (0, 1, 0), (1, 0, 0), 0.8, 0.8)
    sc = native.Scene(ctx, sd)
    got, _ = sc.render_samples(cam, abi.make_params(16, 16, 4, max_path_len=M, min_path_len=M))
    series = rho * np.array(I) / (math.pi * R * R) * (1 - rho ** (M + 1)) / (1 - rho)
    np.testing.assert_allclose(got[:, 2:5], np.tile(series, (len(got), 1)), rtol=2e-5)
    sc.close()


def test_point_light_photon_maps_and_render(ctx):
    sd, cam = lit_cornell()
    pp = abi.photon_params(caustic=2000, direct=20000, indirect=20000, seed=6)
    sc = native.Scene(ctx, sd)
    gm = sc.photon_map(pp)
    _compare_maps(gm, oracle.OracleScene(sd).photon_map(pp))
    arrays = [gm.export(s) for s in range(3)]
    om = oracle.OraclePhotonMap.import_arrays(pp, arrays, gm.info()["paths"])
    p = abi.make_params(24, 20, 4, photon_map=gm, seed=5)
    got, _ = sc.render_samples(cam, p)
    _, ref, _ = oracle.OracleScene(sd).render(cam, abi.make_params(24, 20, 4, photon_map=om, seed=5),
                                              samples=True)
    check_paths(got, ref, 1e-3, 0.95, "point lights, photon integrator")
    gm.close()
    sc.close()


def test_point_light_job_in_batch_equals_alone(ctx):
    sd, cam = lit_cornell()
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
