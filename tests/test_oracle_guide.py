"""Pins the FP64 restatement of the learned proposal (oracle/guide_ref.py):
vMF normalisation and sampling against closed forms, the mixture density
integrates to one, and the analytic gradient equals central finite
differences of the FP64 loss."""
import math

import numpy as np
import pytest

from oracle import guide_ref as G


def sphere_grid(n=400):
    """Midpoint quadrature over (cos theta, phi): weights sum to 4 pi."""
    ct = -1 + (np.arange(n) + 0.5) * 2 / n
    ph = (np.arange(2 * n) + 0.5) * math.pi / n
    CT, PH = np.meshgrid(ct, ph, indexing="ij")
    st = np.sqrt(1 - CT ** 2)
    d = np.stack([st * np.cos(PH), st * np.sin(PH), CT], -1).reshape(-1, 3)
    return d, np.full(len(d), 4 * math.pi / len(d))


@pytest.mark.parametrize("kappa", [1e-6, 0.05, 1.0, 10.0, 100.0, 400.0])
def test_vmf_normalisation(kappa):
    d, wq = sphere_grid(800 if kappa > 50 else 300)
    mu = np.array([0.3, -0.5, 0.81])
    mu /= np.linalg.norm(mu)
    p = G.vmf_pdf(kappa, d @ mu)
    assert (p * wq).sum() == pytest.approx(1.0, abs=2e-3)
    # closed form at the mode: kappa / (2 pi (1 - e^{-2 kappa}))
    if kappa > 1e-3:
        assert G.vmf_pdf(kappa, 1.0) == pytest.approx(kappa / (2 * math.pi * (1 - math.exp(-2 * kappa))))


@pytest.mark.parametrize("kappa", [0.5, 5.0, 80.0])
def test_vmf_sampling_matches_density(kappa):
    """CDF of c = mu . w under the vMF: (e^{kappa (t-1)} - e^{-2 kappa}) / (1 - e^{-2 kappa});
    samples are unit vectors, azimuthally symmetric about mu."""
    rng = np.random.default_rng(7)
    mu = np.array([0.0, 0.6, 0.8])
    s = G.vmf_sample(mu, kappa, rng.random(20000), rng.random(20000))
    assert np.allclose(np.linalg.norm(s, axis=1), 1, atol=1e-9)
    c = np.sort(s @ mu)
    cdf = (np.exp(kappa * (c - 1)) - math.exp(-2 * kappa)) / (1 - math.exp(-2 * kappa))
    ks = np.max(np.abs(cdf - (np.arange(len(c)) + 0.5) / len(c)))
    assert ks < 1.63 / math.sqrt(len(c))
    perp = s - (s @ mu)[:, None] * mu
    assert np.linalg.norm(perp.mean(0)) < 0.02


def random_params(seed, scale=0.3):
    rng = np.random.default_rng(seed)
    P = {}
    for k, s in G.SHAPES:
        P[k] = rng.normal(0, scale, s)
    P["W1"][:, 78:] = 0
    P["W4"][40:] = 0
    P["b4"][40:] = 0
    return G.pack(P)


def random_records(n, seed):
    rng = np.random.default_rng(seed)
    rec = np.zeros((n, 16))
    rec[:, :3] = rng.random((n, 3))
    d = rng.normal(size=(n, 3))
    rec[:, 3:6] = d / np.linalg.norm(d, axis=1, keepdims=True)
    wi = rng.normal(size=(n, 3))
    rec[:, 6:9] = wi / np.linalg.norm(wi, axis=1, keepdims=True)
    rec[:, 9] = rng.exponential(1.0, n) * (rng.random(n) < 0.8)
    return rec


def test_mixture_integrates_to_one():
    w = random_params(1)
    raw, _, _ = G.forward(w, np.array([[0.2, 0.5, 0.7, 0.0, 1.0, 0.0]]))
    d, wq = sphere_grid(600)
    assert (G.mix_pdf(np.repeat(raw, len(d), 0), d) * wq).sum() == pytest.approx(1.0, abs=3e-3)


def test_gradient_matches_finite_differences():
    w = random_params(2)
    rec = random_records(24, 3)
    g, ls, ws = G.gradient(w, rec)
    rng = np.random.default_rng(4)
    idx = np.concatenate([rng.choice(np.arange(G.NPARAM), 60, replace=False),
                          [5120 + 3, 16576 + 7, 13504 + 64 * 39 + 5]])  # a bias of each kind
    h = 1e-6
    for i in idx:
        if np.all(w == w):
            wp, wm = w.copy(), w.copy()
            wp[i] += h
            wm[i] -= h
            fd = (G.loss_sum(wp, rec)[0] - G.loss_sum(wm, rec)[0]) / (2 * h) / ws
            assert g[i] == pytest.approx(fd, rel=1e-5, abs=1e-8), i
    # padding parameters get no gradient
    P = G.unpack(g)
    assert np.all(P["W1"][:, 78:] == 0) and np.all(P["W4"][40:] == 0) and np.all(P["b4"][40:] == 0)


def test_raw_gradient_matches_finite_differences():
    rng = np.random.default_rng(5)
    raw = rng.normal(0, 1, (6, 40))
    raw[:, 3::5] = rng.uniform(-2, 5, (6, 8))  # inside the concentration clamp
    wi = rng.normal(size=(6, 3))
    wi /= np.linalg.norm(wi, axis=1, keepdims=True)
    g = G.raw_grad(raw, wi)
    h = 1e-6
    for k in range(40):
        rp, rm = raw.copy(), raw.copy()
        rp[:, k] += h
        rm[:, k] -= h
        fd = (-np.log(G.mix_pdf(rp, wi)) + np.log(G.mix_pdf(rm, wi))) / (2 * h)
        assert np.allclose(g[:, k], fd, rtol=1e-5, atol=1e-8), k
