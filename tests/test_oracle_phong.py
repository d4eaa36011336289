"""Phong on the oracle (material/phong.cc:38-153, material/phong-dist.h):
a diffuse layer chosen with probability intensity(colour) and a Phong
lobe of the half vector, pdf (e + 1) / (2 pi) cos^e.  Pinned by the
sample/density consistency (which fixes the lobe's normalisation, the
1 / (4 v.h) Jacobian and the layer weights), by the directional albedo
of the lobe against a one-dimensional quadrature, and by the diffuse
branch's cos(theta) = sqrt(u / w)."""
import math

import numpy as np
import pytest
from scipy import integrate

import oracle
from snogray_amd import abi
from tests.test_oracle_shading import ALL_NONSPEC, grid_params, uniform_hemisphere, unit

PHONG = abi.SGR_MAT_PHONG
SPECULAR, REFLECTIVE = 0x01, 0x10


@pytest.mark.parametrize("e,diffuse", [(8.0, 0.0), (40.0, 0.0), (20.0, 0.3), (3.0, 0.6)])
def test_phong_density_matches_sampling(e, diffuse):
    """The fraction of parameter space giving a valid direction equals the
    eval pdf's integral over the hemisphere (uniform Monte Carlo), and
    eval at a sampled direction gives the sampled value and pdf."""
    mat = {"type": PHONG, "color": (diffuse,) * 3, "gloss_color": (0.5, 0.4, 0.3), "m": e}
    v, gn = unit((0.4, -0.2, 0.85)), (0, 0, 1)
    s, _ = oracle.bsdf_probe(mat, v, gn, grid_params(400), np.zeros((0, 3)),
                             flags_s=ALL_NONSPEC)
    ok = s[:, 3] > 0
    dirs = uniform_hemisphere(400000, 3)
    _, ev = oracle.bsdf_probe(mat, v, gn, np.zeros((0, 2)), dirs, flags_e=ALL_NONSPEC)
    x = ev[:, 3] * 2 * np.pi
    se = x.std() / math.sqrt(len(x))
    assert abs(x.mean() - ok.mean()) < 4 * se + 2e-3
    d = s[ok][::97, 4:7]
    _, e2 = oracle.bsdf_probe(mat, v, gn, np.zeros((0, 2)), d, flags_e=ALL_NONSPEC)
    assert np.allclose(e2, s[ok][::97, :4], rtol=2e-3, atol=1e-6)
    # every sample is tagged REFLECTIVE only (not DIFFUSE or GLOSSY)
    assert np.all(s[ok, 7].astype(int) == REFLECTIVE)


@pytest.mark.parametrize("e", [5.0, 30.0])
def test_phong_albedo_at_normal_incidence(e):
    """View along the normal, specular layer only: the importance-sampled
    albedo E[f cos / pdf] equals ks (e + 1) int_0^{pi/2} cos^e(t / 2)
    cos t sin t dt (h bisects v = n and l, so n.h = cos(t / 2))."""
    ks = np.array([0.5, 0.4, 0.3])
    mat = {"type": PHONG, "color": (0, 0, 0), "gloss_color": tuple(ks), "m": e}
    s, _ = oracle.bsdf_probe(mat, (0, 0, 1), (0, 0, 1), grid_params(600), np.zeros((0, 3)),
                             flags_s=ALL_NONSPEC)
    ok = s[:, 3] > 0
    est = np.where(ok[:, None], s[:, :3] * s[:, 6:7] / np.where(ok, s[:, 3], 1)[:, None],
                   0).mean(0)
    q, _ = integrate.quad(lambda t: math.cos(t / 2) ** e * math.cos(t) * math.sin(t), 0,
                          math.pi / 2)
    assert np.allclose(est, ks * (e + 1) * q, rtol=3e-3)


def test_phong_diffuse_branch_and_specular_flags():
    """u below w = intensity(colour) picks the cosine lobe with u / w:
    cos(theta) = sqrt(u / w); the value then holds colour / pi plus the
    Phong term of the half vector.  With only specular layers asked for,
    the value and pdf are zero (nothing to recurse into)."""
    col = (0.2, 0.3, 0.4)      # intensity 0.3
    mat = {"type": PHONG, "color": col, "gloss_color": (0.5, 0.5, 0.5), "m": 10.0}
    prm = grid_params(100)
    s, _ = oracle.bsdf_probe(mat, (0, 0, 1), (0, 0, 1), prm, np.zeros((0, 3)),
                             flags_s=ALL_NONSPEC)
    lo = prm[:, 0] < 0.3 * (1 - 1e-6)
    assert np.allclose(s[lo, 6], np.sqrt(prm[lo, 0] / 0.3), rtol=1e-5)
    l = s[lo, 4:7]
    h = l + np.array([0, 0, 1.0])
    h /= np.linalg.norm(h, axis=1, keepdims=True)
    spec = 11.0 / (2 * np.pi) * h[:, 2] ** 10
    assert np.allclose(s[lo, :3], np.array(col) / np.pi + 0.5 * spec[:, None], rtol=1e-4)
    s2, _ = oracle.bsdf_probe(mat, (0, 0, 1), (0, 0, 1), prm, np.zeros((0, 3)),
                              flags_s=SPECULAR | REFLECTIVE)
    assert np.all(s2[:, :4] == 0)
