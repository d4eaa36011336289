"""Stencil on the oracle (material/stencil.cc): the underlying BSDF with
probability intensity(opacity), its value times the opacity and its pdf
times 1 / intensity; else straight through with value (1 - opacity) /
cos and the remaining probability; shadow rays pass it with
transmittance max(1 - opacity, 0).  Pinned by a floor lit through a
stencil pane (Lambert's polygon formula times 1 - opacity), a stencil
pane in front of a uniform sky (both branches in expectation), and a
stencil floor under a quad light, whose direct lighting mixes the
stencil's scaled value and pdf (light samples) with the bare underlying
sample (BSDF samples, which ask for no specular part) -- by quadrature."""
import math

import numpy as np
import pytest

import oracle
from snogray_amd import abi, scenes
from tests.test_oracle_pins import IDENTITY_FRAME, camera_dirs
from tests.test_oracle_shading import _floor_points, _polygon_irradiance, unit

LE = 8.0
V0, E1, E2 = np.array([-0.6, 2.0, -0.4]), np.array([0.0, 0.0, 0.9]), np.array([1.1, 0, 0])
QUAD = np.array([V0, V0 + E1, V0 + E1 + E2, V0 + E2])


def test_floor_lit_through_a_stencil_pane():
    op = np.array([0.3, 0.5, 0.8])
    sd = scenes.SceneDesc()
    sd.tripar((-50, 0, -50), (100, 0, 0), (0, 0, 100), sd.lambert((0.5, 0.5, 0.5)),
              parallelogram=True)
    sd.tripar((-5, 1.2, -5), (10, 0, 0), (0, 0, 10),
              sd.stencil(tuple(op), sd.lambert((0.2, 0.2, 0.2))), parallelogram=True)
    sd.tripar(tuple(V0), tuple(E1), tuple(E2), sd.glow((LE, LE, LE)), parallelogram=True)
    cam = scenes.camera((0.0, 0.8, 0.0), (0, -1, 0), (0, 0, 1), (1, 0, 0), 0.6, 0.6)
    W = 16
    p = abi.make_params(W, W, 64, max_path_len=0, min_path_len=0, seed=5)
    _, smp, _ = oracle.OracleScene(sd).render(cam, p, samples=True)
    pts = _floor_points(cam, smp[:, 0].astype(float), smp[:, 1].astype(float), W, W)
    expect = np.outer(0.5 / np.pi * LE * _polygon_irradiance(pts, QUAD), 1 - op)
    diff = smp[:, 2:5].astype(float) - expect
    se = diff.std(0) / math.sqrt(len(diff))
    assert np.all(np.abs(diff.mean(0)) < 3 * se + 1e-4 * expect.mean(0)), (diff.mean(0), se)


def test_stencil_pane_in_sky_both_branches():
    """direct_samples 0 (the background counts after any bounce): through
    with probability 1 - w (w = intensity(opacity)), weight (1 - o) /
    (1 - w); else the underlying Lambert sample, weight rho o w (value
    rho o / pi, pdf cos / (pi w)), which returns the sky: mean
    E ((1 - o) + rho o w^2)."""
    E, rho = 2.0, 0.6
    op = np.array([0.2, 0.4, 0.9])
    w = op.mean()
    sd = scenes.SceneDesc()
    sd.tripar((-5, -5, 0), (10, 0, 0), (0, 10, 0), sd.stencil(tuple(op), sd.lambert((rho,) * 3)),
              parallelogram=True)
    sd.envmap(np.full((16, 32, 3), E, np.float32), frame=IDENTITY_FRAME)
    cam = scenes.camera((0.3, 0.2, -3), unit((0.1, 0.05, 1)), (0, 1, 0), (1, 0, 0), 0.3, 0.3)
    p = abi.make_params(6, 6, 256, max_path_len=4, min_path_len=10, direct_samples=0, seed=3)
    _, smp, _ = oracle.OracleScene(sd).render(cam, p, samples=True)
    expect = E * ((1 - op) + rho * op * w * w)
    x = smp[:, 2:5].astype(float)
    se = x.std(0) / math.sqrt(len(x))
    assert np.all(np.abs(x.mean(0) - expect) < 3 * se + 1e-4), (x.mean(0), expect, se)
    # a value of the through branch alone
    thru = np.abs(x - E * (1 - op) / (1 - w)).max(1) < 1e-4
    assert 1 - w - 0.05 < thru.mean() < 1 - w + 0.05


def test_stencil_floor_under_a_quad_light():
    """Light samples: value (rho / pi) o, pdf reported (cos / pi) / w in
    the power-heuristic weight; BSDF samples (no specular part asked for)
    come from the underlying Lambert BSDF unscaled.  Expectation per
    floor point, with p_l the light's solid-angle density:
    Le Int cos [(rho o / pi) w_l + (rho / pi) w_b] dw,
    w_l = p_l^2 / (p_l^2 + (cos / (pi w))^2), w_b = (cos / pi)^2 /
    ((cos / pi)^2 + p_l^2)."""
    rho = 0.5
    op = np.array([0.3, 0.6, 0.9])
    w = op.mean()
    sd = scenes.SceneDesc()
    # a large light close to the floor: BSDF samples carry much of the
    # estimate, so the stencil's scaled pdf in the light samples' weights
    # shows
    v0, e1, e2 = np.array([-3.0, 1.2, -3.0]), np.array([0.0, 0, 6.0]), np.array([6.0, 0, 0])
    sd.tripar((-50, 0, -50), (100, 0, 0), (0, 0, 100), sd.stencil(tuple(op), sd.lambert((rho,) * 3)),
              parallelogram=True)
    sd.tripar(tuple(v0), tuple(e1), tuple(e2), sd.glow((LE, LE, LE)), parallelogram=True)
    cam = scenes.camera((0.0, 0.6, 0.0), (0, -1, 0), (0, 0, 1), (1, 0, 0), 0.6, 0.6)
    W = 10
    p = abi.make_params(W, W, 64, max_path_len=0, min_path_len=0, seed=7)
    _, smp, _ = oracle.OracleScene(sd).render(cam, p, samples=True)
    pts = _floor_points(cam, smp[:, 0].astype(float), smp[:, 1].astype(float), W, W)
    nq = np.cross(e2, e1)
    area = np.linalg.norm(nq)
    nq /= area
    g = (np.arange(96) + 0.5) / 96
    a, b = np.meshgrid(g, g, indexing="ij")
    q = v0 + a.reshape(-1, 1) * e1 + b.reshape(-1, 1) * e2

    def radiance(x):
        dv = q - x
        dist = np.linalg.norm(dv, axis=1)
        l = dv / dist[:, None]
        cos_l = np.abs(l @ nq)
        cos = l[:, 1]
        pl = dist ** 2 / (area * cos_l)
        pbr = cos / np.pi / w
        pb = cos / np.pi
        wl = pl ** 2 / (pl ** 2 + pbr ** 2)
        wb = pb ** 2 / (pb ** 2 + pl ** 2)
        # integrand per unit solid angle, converted to area: dw = cos_l dA / d^2
        f = LE * cos[:, None] * (rho * np.outer(wl, op) / np.pi + (rho / np.pi * wb)[:, None])
        return (f * (cos_l / dist ** 2)[:, None]).mean(0) * area

    expect = np.array([radiance(x) for x in pts])
    diff = smp[:, 2:5].astype(float) - expect
    se = diff.std(0) / math.sqrt(len(diff))
    assert np.all(np.abs(diff.mean(0)) < 3 * se + 2e-3 * expect.mean(0)), (diff.mean(0), se)
    # the plain estimator (true pdfs, value rho o / pi) would be rho o E / pi: off
    quad = np.array([v0, v0 + e1, v0 + e1 + e2, v0 + e2])
    plain = np.outer(rho / np.pi * LE * _polygon_irradiance(pts, quad), op)
    assert np.any(np.abs(plain.mean(0) - smp[:, 2:5].mean(0)) > 10 * se)


def test_stencil_nesting_and_models_refused():
    sd = scenes.SceneDesc()
    s1 = sd.stencil((0.5, 0.5, 0.5), sd.lambert((0.5, 0.5, 0.5)))
    sd.sphere((0, 0, 0), 1.0, sd.stencil((0.5, 0.5, 0.5), s1))
    with pytest.raises(RuntimeError, match="stencil"):
        oracle.OracleScene(sd)


def _opaque_pair(under):
    """The same room twice: its ball once of a fully opaque stencil over
    UNDER, once of UNDER itself."""
    out = []
    for wrap in (True, False):
        sd = scenes.SceneDesc()
        sd.tripar((-50, 0, -50), (100, 0, 0), (0, 0, 100), sd.lambert((0.6, 0.6, 0.6)),
                  parallelogram=True)
        sd.tripar(V0, E1, E2, sd.glow((LE,) * 3), parallelogram=True)
        m = under(sd)
        sd.sphere((0.0, 0.8, 0.0), 0.8, sd.stencil((1.0, 1.0, 1.0), m) if wrap else m)
        # thin glass reflects by the medium it is met in (air or the ball's)
        sd.sphere((0.0, 0.8, 0.0), 0.4, sd.thin_glass((1.0, 1.0, 1.0), 1.3))
        out.append(sd)
    return out


@pytest.mark.parametrize("under", [
    lambda sd: sd.lambert((0.7, 0.3, 0.5)),
    lambda sd: sd.cook_torrance((0.3, 0.2, 0.1), (0.5, 0.5, 0.5), 0.3, 1.7),
    lambda sd: sd.glass(1.5),
    lambda sd: sd.glow((2.0, 2.0, 2.0)),   # no BSDF
])
def test_fully_opaque_stencil_is_its_underlying_bsdf(under):
    """material/stencil.cc:140-148: opacity >= 1 hands over the
    underlying material's BSDF (none for a bare glow), with the stencil's
    own medium (none) and emission (none).  Over Lambert or Cook-Torrance
    every sample equals the bare material's; over glass the ball keeps no
    medium, so a thin glass ball inside it reflects as in air and the
    images differ; over a bare glow a camera ray that
    meets the ball ends there with nothing, where the glow itself gives
    its colour."""
    a, b = _opaque_pair(under)
    W, H = 24, 16
    cam = scenes.look_at((0, 1.6, -5), (0, 0.8, 0), vfov_deg=40.0)
    p = abi.make_params(W, H, 4, max_path_len=6, direct_samples=1)
    _, sa, _ = oracle.OracleScene(a).render(cam, p, samples=True)
    _, sb, _ = oracle.OracleScene(b).render(cam, p, samples=True)
    probe = scenes.SceneDesc()
    t = probe.materials[under(probe)]["type"]
    if t == abi.SGR_MAT_GLASS:
        assert not np.array_equal(sa[:, 2:5], sb[:, 2:5])
    elif t == abi.SGR_MAT_GLOW:
        d = camera_dirs(cam, sa[:, 0].astype(float), sa[:, 1].astype(float), W, H)
        o = np.array([0.0, 1.6, -5.0]) - np.array([0.0, 0.8, 0.0])
        bq = d @ o
        disc = bq * bq - (d * d).sum(1) * (o @ o - 0.64)
        ball = disc > 0.05 * (d * d).sum(1)   # clear of the silhouette
        assert ball.sum() > 20
        assert np.all(sa[ball, 2:5] == 0.0)
        assert np.all(sb[ball, 2:5] == 2.0)
    else:
        np.testing.assert_array_equal(sa, sb)
