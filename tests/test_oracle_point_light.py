"""The oracle's point lights (light/point-light.cc, light/point-light.h)
against closed forms: the light sample's value, direction and distance
from points facing toward and away from the light, a spotlight's cone and
fringe ramp, the zero eval, a point light at the centre of a diffuse
sphere (every camera sample's radiance is the finite series of its
bounces, which fails if the light sample were weighted by multiple
importance or the 1/d^2 fall-off dropped), and the photons' flux
(4 pi I for an omni light, the cone's integral for a spotlight)."""
import math

import numpy as np
import pytest

import oracle
from snogray_amd import abi, scenes
from tests.test_oracle_photon import _enclosed, _flux

I = (3.0, 2.0, 1.0)


def _lit_floor(**kw):
    sd = scenes.SceneDesc()
    sd.tripar((-5, 0, -5), (10, 0, 0), (0, 0, 10), sd.lambert((0.5, 0.5, 0.5)), parallelogram=True)
    sd.point_light((0.25, 1.5, -0.375), I, **kw)   # exact in float32
    return sd


@pytest.mark.parametrize("fp64", [True, False])
def test_point_light_sample(fp64):
    sd = _lit_floor()
    sc = oracle.OracleScene(sd, fp64=fp64)
    rng = np.random.default_rng(3)
    pts = np.c_[rng.uniform(-2, 2, 50), rng.uniform(-0.5, 1.0, 50), rng.uniform(-2, 2, 50)]
    nrm = rng.normal(size=(50, 3))
    nrm /= np.linalg.norm(nrm, axis=1)[:, None]
    out = sc.light_probe(0, pts, nrm, uv=rng.uniform(size=(50, 2)))
    lp = np.array([0.25, 1.5, -0.375])
    d = lp - pts
    dist = np.linalg.norm(d, axis=1)
    w = d / dist[:, None]
    facing = np.sum(w * nrm, 1) > 0
    tol = 1e-12 if fp64 else 2e-6
    assert facing.sum() > 10 and (~facing).sum() > 10
    np.testing.assert_allclose(out["val"][facing], np.outer(1 / dist[facing] ** 2, I), rtol=tol)
    np.testing.assert_array_equal(out["pdf"][facing], 1.0)
    np.testing.assert_allclose(out["dist"][facing], dist[facing], rtol=tol)
    np.testing.assert_allclose(out["dir"][facing], w[facing], atol=10 * tol)
    assert np.all(out["pdf"][~facing] == 0) and np.all(out["val"][~facing] == 0)
    # no direction meets a point
    ev = sc.light_probe(0, pts, nrm, dirs=w)
    assert np.all(ev["pdf"] == 0) and np.all(ev["val"] == 0)


@pytest.mark.parametrize("dlen", [1.0, 0.5])
def test_spotlight_cone_and_fringe(dlen):
    """A spotlight pointing down: full intensity inside the core,
    ((cos - cos_half) / (cos_core - cos_half))^4 across the fringe, nothing
    outside the cone.  The cosine is taken against the direction as given
    (PointLight keeps it as its frame's z axis unnormalised), so a
    direction of length 1/2 halves it."""
    ang, fr = 1.2, 0.4
    sd = scenes.SceneDesc()
    sd.tripar((-5, 0, -5), (10, 0, 0), (0, 0, 10), sd.lambert((0.5, 0.5, 0.5)), parallelogram=True)
    lp = np.array([0.0, 2.0, 0.0])
    sd.point_light(tuple(lp), I, angle=ang, dir=(0.0, -dlen, 0.0), fringe_angle=fr)
    sc = oracle.OracleScene(sd, fp64=True)
    x = np.linspace(0, 4.0, 200)
    pts = np.c_[x, np.zeros_like(x), np.zeros_like(x)]
    out = sc.light_probe(0, pts, np.tile([0.0, 1.0, 0.0], (len(x), 1)), uv=np.zeros((len(x), 2)))
    dist = np.linalg.norm(lp - pts, axis=1)
    cos_dir = 2.0 * dlen / dist          # ((p - l) / |p - l|) . (0, -dlen, 0)
    a32, f32 = float(np.float32(ang)), float(np.float32(fr))   # the descriptor's floats
    ch, cc = math.cos(a32 / 2), math.cos((a32 - f32) / 2)
    f = np.where(cos_dir < cc, ((cos_dir - ch) / (cc - ch)) ** 4, 1.0)
    f = np.where(cos_dir >= ch, f, 0.0)
    np.testing.assert_allclose(out["val"], np.outer(f / dist ** 2, I), rtol=1e-12, atol=1e-300)
    inside = cos_dir >= ch
    if dlen == 1.0:
        assert inside.sum() > 20 and (~inside).sum() > 20 and np.sum(inside & (cos_dir < cc)) > 5
    else:
        assert not inside.any()
    assert np.all(out["pdf"][inside] == 1) and np.all(out["pdf"][~inside] == 0)


@pytest.mark.parametrize("fp64", [True, False])
def test_point_light_in_a_diffuse_sphere(fp64):
    """A point light of intensity I at the centre of a Lambert sphere
    (radius R, albedo rho), seen from inside: every vertex receives I / R^2
    at normal incidence and the cosine-sampled bounce carries throughput
    rho, so without Russian roulette (min path length = max = M) each
    camera sample is rho I / (pi R^2) (1 + rho + ... + rho^M) exactly; the
    direct-lighting integrator gives the first term."""
    R, rho, M = 2.0, 0.5, 20
    sd = scenes.SceneDesc()
    sd.sphere((0.0, 0.0, 0.0), R, sd.lambert((rho,) * 3))
    sd.point_light((0.0, 0.0, 0.0), I)
    cam = scenes.camera((0.3, 0.1, -0.2), (0, 0, 1), (0, 1, 0), (1, 0, 0), 0.8, 0.8)
    sc = oracle.OracleScene(sd, fp64=fp64)
    tol = 1e-7 if fp64 else 2e-5   # Lambert's 1/pi is the reference's float INV_PIf
    first = rho * np.array(I) / (math.pi * R * R)
    p = abi.make_params(6, 6, 2, max_path_len=M, min_path_len=M)
    _, smp, _ = sc.render(cam, p, samples=True)
    series = first * (1 - rho ** (M + 1)) / (1 - rho)
    np.testing.assert_allclose(smp[:, 2:5], np.tile(series, (len(smp), 1)), rtol=tol)
    _, smp, _ = sc.render(cam, abi.make_params(6, 6, 2, integrator=abi.SGR_INTEG_DIRECT),
                          samples=True)
    np.testing.assert_allclose(smp[:, 2:5], np.tile(first, (len(smp), 1)), rtol=tol)


@pytest.mark.parametrize("kind", ["omni", "spot", "fringe"])
def test_point_light_photon_flux(kind):
    """Photons from a point light inside a black box: the first-hit power
    per path is the light's flux, 4 pi I in all directions, 2 pi (1 -
    cos(angle/2)) I for a cone, and with a fringe the d^4 ramp adds
    2 pi (cos_core - cos_half) / 5 I over the fringe's band."""
    ang, fr = {"omni": (0.0, 0.0), "spot": (1.4, 0.0), "fringe": (1.4, 0.6)}[kind]

    def add(sd):
        sd.point_light((0.1, 0.2, -0.1), I, angle=ang, dir=(0.3, -1.0, 0.2), fringe_angle=fr)
    flux, info = _flux(_enclosed(add))
    assert info["paths"][0] == info["photons"][0]
    if kind == "omni":
        phi = 4 * math.pi
    else:
        ch, cc = math.cos(ang / 2), math.cos((ang - fr) / 2)
        phi = 2 * math.pi * ((1 - cc) + (cc - ch) / 5)
    np.testing.assert_allclose(flux, phi * np.array(I), rtol=2e-3)
