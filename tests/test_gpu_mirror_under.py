"""Mirrors over a material (material/mirror.h:37-42) on the device against
the oracle: per-path radiance of the path tracer (reflectance 0, where
every sample is the underlying BSDF's, and reflectance 0.6 over tilted
vertex normals, where back-facing views fall through to it), the direct
integrator, photon maps (a mirror over a diffuse material holds photons)
and the refusal of an unsupported underlying material."""
import numpy as np
import pytest

import oracle
from snogray_amd import abi, native, scenes
from tests.test_gpu_photon import _compare_maps
from tests.test_gpu_render import check_paths

pytestmark = pytest.mark.gpu


def mirror_scene(refl):
    sd, cam = scenes.cornell_box()
    ct = sd.cook_torrance((0.4, 0.3, 0.3), (0.5, 0.5, 0.5), 0.2, 1.7)
    ph = sd.phong((0.3, 0.4, 0.3), (0.4, 0.4, 0.4), 30.0)
    la = sd.lambert((0.6, 0.5, 0.4))
    sd.sphere((-1.2, 0.8, 0.5), 0.8, sd.mirror((refl,) * 3, underlying=ct))
    sd.sphere((1.3, 0.7, -0.3), 0.7, sd.mirror((refl,) * 3, underlying=ph))
    v, t = scenes.icosphere(2)
    v = np.asarray(v, np.float32) * 0.6 + np.array([0.0, 2.4, 0.8], np.float32)
    n = (v - np.array([0.0, 2.4, 0.8], np.float32)) + np.float32(0.3)   # tilted normals
    sd.mesh(v, t, sd.mirror((refl,) * 3, underlying=la), left_handed=False, normals=n)
    return sd, cam


@pytest.mark.parametrize("refl", [0.0, 0.6])
def test_mirror_over_material_paths_match_oracle(ctx, refl):
    sd, cam = mirror_scene(refl)
    sc, osc = native.Scene(ctx, sd), oracle.OracleScene(sd)
    p = abi.make_params(40, 32, 8, max_path_len=8, seed=7)
    got, gst = sc.render_samples(cam, p)
    _, ref, ost = osc.render(cam, p, samples=True)
    check_paths(got, ref, 1e-4, 0.99, f"mirror over materials, reflectance {refl}")
    for k in ("scene_intersect_calls", "shadow_tests", "illum_calls"):
        assert abs(gst[k] - ost[k]) <= 0.002 * ost[k] + 2, (k, gst[k], ost[k])
    q = abi.make_params(40, 32, 4, integrator=abi.SGR_INTEG_DIRECT, light_samples=2, seed=7)
    got, _ = sc.render_samples(cam, q)
    _, ref, _ = osc.render(cam, q, samples=True)
    check_paths(got, ref, 1e-4, 0.99, f"mirror over materials, direct integrator, {refl}")
    sc.close()


def test_mirror_over_material_photon_maps(ctx):
    sd, _ = mirror_scene(0.0)
    pp = abi.photon_params(caustic=2000, direct=20000, indirect=20000, seed=2)
    sc = native.Scene(ctx, sd)
    _compare_maps(sc.photon_map(pp), oracle.OracleScene(sd).photon_map(pp))
    sc.close()


def test_mirror_underlying_refused(ctx):
    sd = scenes.SceneDesc()
    g = sd.glass(1.5)
    sd.tripar((0, 0, 0), (1, 0, 0), (0, 0, 1), sd.mirror((0.5, 0.5, 0.5), underlying=g))
    with pytest.raises(native.SgrError, match="lambert, cook-torrance or phong"):
        native.Scene(ctx, sd)
