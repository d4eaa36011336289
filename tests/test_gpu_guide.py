"""Learned proposal on the GPU: the MFMA forward pass and the training
gradient against the CPU FP64 restatement, the first Adam step's closed
form, training on a known target distribution, unbiasedness of guided
rendering, delta-BSDF bypass, and what a trained guide gains at equal
samples per pixel."""
import math

import numpy as np
import pytest

from oracle import guide_ref as G
from snogray_amd import abi, native, scenes
from tests.test_oracle_guide import random_params, random_records

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def guide(ctx):
    g = native.Guide(ctx, seed=3)
    yield g
    g.close()


def test_initial_parameters(guide):
    w = guide.params()
    assert len(w) == G.NPARAM == abi.GD_NPARAM
    P = G.unpack(w)
    assert np.all(P["W1"][:, 78:] == 0) and np.all(P["W4"][40:] == 0)
    assert np.all(P["b1"] == 0) and np.all(P["b4"] == 0)
    assert 0.5 < P["W2"].std() * math.sqrt(64 / 2) < 1.5   # He initialisation


def test_forward_matches_cpu(guide):
    w = random_params(11).astype(np.float32)
    guide.set_params(w)
    rng = np.random.default_rng(12)
    x = np.concatenate([rng.random((3000, 3)), rng.normal(size=(3000, 3))], 1)
    x[:, 3:] /= np.linalg.norm(x[:, 3:], axis=1, keepdims=True)
    out = guide.forward(x.astype(np.float32))
    ref_e, _, _ = G.forward(w, x.astype(np.float32), emulate=True)
    err = np.abs(out - ref_e)
    assert np.median(err / (np.abs(ref_e) + 1e-2)) < 1e-3
    assert np.mean(err <= 1e-2 * np.abs(ref_e) + 1e-2) > 0.995
    ref, _, _ = G.forward(w, x)           # full FP64, no bf16 roundings
    assert np.linalg.norm(out - ref) / np.linalg.norm(ref) < 2e-2
    L = guide.forward(x.astype(np.float32), lobes=True).reshape(-1, 8, 5)
    assert np.allclose(np.linalg.norm(L[..., :3], axis=-1), 1, atol=1e-5)
    assert np.allclose(L[..., 4].sum(-1), 1, atol=1e-5)
    assert np.all((L[..., 3] >= math.exp(G.KLO) * 0.999) & (L[..., 3] <= math.exp(G.KHI) * 1.001))
    assert np.allclose(L, G.lobes(out).astype(np.float32), rtol=1e-4, atol=1e-5)


def test_gradient_matches_cpu_fp64(guide):
    w = random_params(21).astype(np.float32)
    guide.set_params(w)
    rec = random_records(6000, 22).astype(np.float32)
    g, (ls, ws) = guide.train_batch(rec, lr=0.0)
    ge, lse, wse = G.gradient(w, rec, emulate=True)
    assert ws == pytest.approx(wse, rel=1e-5)
    assert ls == pytest.approx(lse, rel=2e-3)
    rel = np.linalg.norm(g - ge) / np.linalg.norm(ge)
    cos = g @ ge / (np.linalg.norm(g) * np.linalg.norm(ge))
    assert rel < 2e-2 and cos > 0.9998, (rel, cos)
    g64, _, _ = G.gradient(w, rec)        # full FP64, no bf16 roundings
    rel64 = np.linalg.norm(g - g64) / np.linalg.norm(g64)
    assert rel64 < 6e-2, rel64
    # per layer, so no layer's error hides behind another's magnitude
    for k in ("W1", "W2", "W3", "W4", "b1", "b4"):
        a, b = G.unpack(g)[k], G.unpack(ge)[k]
        assert np.linalg.norm(a - b) <= 3e-2 * np.linalg.norm(b) + 1e-7, k
    # zero-weight records contribute nothing
    rec0 = rec.copy()
    rec0[::2, 9] = 0
    g0, _ = guide.train_batch(rec0, lr=0.0)
    ge0, _, _ = G.gradient(w, rec0, emulate=True)
    assert np.linalg.norm(g0 - ge0) / np.linalg.norm(ge0) < 2e-2


def test_first_adam_step_closed_form(guide):
    w = random_params(31).astype(np.float32)
    guide.set_params(w)
    rec = random_records(2000, 32).astype(np.float32)
    g, _ = guide.train_batch(rec, lr=1e-3)
    w1 = guide.params()
    # bias-corrected first step: m_hat = g, v_hat = g^2
    expect = w - 1e-3 * g / (np.abs(g) + 1e-8)
    assert np.allclose(w1, expect, rtol=1e-6, atol=1e-7)


def test_training_fits_a_known_distribution(guide):
    """Incoming directions drawn from a vMF (kappa 20) about +y: the
    weighted -log q falls from the broad initial lobes to near the
    target's entropy, and the learned density peaks at +y."""
    guide.set_params(np.asarray(native.Guide(guide.ctx, seed=5).params()))
    rng = np.random.default_rng(41)
    mu, kappa = np.array([0.0, 1.0, 0.0]), 20.0
    losses = []
    for step in range(400):
        n = 4096
        rec = np.zeros((n, 16), np.float32)
        rec[:, :3] = rng.random((n, 3))
        d = rng.normal(size=(n, 3))
        rec[:, 3:6] = d / np.linalg.norm(d, axis=1, keepdims=True)
        rec[:, 6:9] = G.vmf_sample(mu, kappa, rng.random(n), rng.random(n))
        rec[:, 9] = 1.0
        _, (ls, ws) = guide.train_batch(rec, lr=1e-3)
        losses.append(ls / ws)
    # entropy of vMF(20): -(log(kappa / (2 pi (1 - e^-2k))) - 1 + 1/kappa... ) ~ -0.11
    ent = -(math.log(kappa / (2 * math.pi * (1 - math.exp(-2 * kappa)))) - 1 + 1 / kappa)
    assert losses[0] > 2.0                  # near the uniform density's log(4 pi) = 2.53
    assert np.mean(losses[-20:]) < ent + 0.35
    x = np.array([[0.5, 0.5, 0.5, 0.0, 0.0, 1.0]], np.float32)
    raw = guide.forward(x)
    assert G.mix_pdf(raw, [mu])[0] > 5 * G.mix_pdf(raw, [[0.0, -1.0, 0.0]])[0]


def test_guided_render_is_unbiased(ctx):
    """Cornell box: per-pixel means over independent seeds with a trained
    guide (fixed) and with plain BSDF sampling agree within 3 standard
    errors."""
    sd, cam = scenes.cornell_box()
    sc = native.Scene(ctx, sd)
    g = native.Guide(ctx, seed=7)
    for s in range(4):  # online training
        sc.render_guided(cam, abi.make_params(48, 48, 16, max_path_len=8, seed=100 + s), g)
    assert g.stats()["train_steps"] > 0
    K = 24

    def imgs(guided):
        out = []
        for s in range(K):
            p = abi.make_params(24, 24, 16, max_path_len=8, seed=1000 + s + (500 if guided else 0))
            a = sc.render_guided(cam, p, g, train=False)[0] if guided else sc.render(cam, p)[0]
            out.append(native.finalize(a)[..., :3])
        return np.stack(out)

    a, b = imgs(True), imgs(False)
    ma, mb = a.mean(0), b.mean(0)
    se = np.sqrt(a.var(0, ddof=1) / K + b.var(0, ddof=1) / K)
    frac = np.mean(np.abs(ma - mb) <= 3 * se + 1e-6)
    assert frac > 0.97, frac
    tot_a, tot_b = a.mean((1, 2, 3)), b.mean((1, 2, 3))
    assert abs(tot_a.mean() - tot_b.mean()) < 3 * math.sqrt(tot_a.var(ddof=1) / K +
                                                            tot_b.var(ddof=1) / K)
    g.close()
    sc.close()


def test_delta_bsdfs_bypass_the_guide(ctx):
    """Only mirror and glass surfaces under an envmap: no vertex is
    guided, so the guided image equals the plain one bit for bit."""
    sd = scenes.SceneDesc()
    sd.sphere((0, 0, 0), 1.0, sd.mirror((0.9, 0.8, 0.7)))
    sd.sphere((2.2, 0, 0), 0.8, sd.glass(1.5))
    sd.envmap(scenes.gaussian_lobe_envmap(64, 32, 3))
    cam = scenes.look_at((1, 0.5, -5), (1, 0, 0), vfov_deg=50)
    sc = native.Scene(ctx, sd)
    g = native.Guide(ctx, seed=9)
    p = abi.make_params(32, 32, 4, max_path_len=6, seed=5)
    a, _ = sc.render_guided(cam, p, g, train=True)
    b, _ = sc.render(cam, p)
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    g.close()
    sc.close()


def test_trained_guide_gain_at_equal_spp(ctx):
    """relMSE against a 2048-spp plain render of guided renders (the
    guide trained online on the same view, four 512 x 512 renders of 16
    spp as in a benchmark frame) and of plain renders, both at 16 spp, on
    a fixed set of 8 seeds (renders and training are deterministic): the
    median of the per-seed ratios must not fall below 0.8.  Per seed the
    ratio is decided by whether the plain render holds one of the Russian
    roulette's rare, very bright samples (4 of these 8 do: median 3.8);
    in pixels without one the two estimators are about equal per sample,
    so this asserts that the guide does no harm, not the equal-time
    target, whose measurement needs the 1080p frame's many pixels."""
    sd, cam = scenes.cornell_box()
    sc = native.Scene(ctx, sd)
    W = 96
    ref = native.finalize(sc.render(cam, abi.make_params(W, W, 2048, max_path_len=8,
                                                         seed=77))[0])[..., :3]
    g = native.Guide(ctx, seed=8)
    for s in range(4):
        sc.render_guided(cam, abi.make_params(512, 512, 16, max_path_len=8, seed=200 + s), g)

    def relmse(x):
        return float(np.mean((x - ref) ** 2 / (ref ** 2 + 1e-2)))
    ratios, rg, rp = [], [], []
    for s in range(8):
        p = abi.make_params(W, W, 16, max_path_len=8, seed=300 + s)
        rg.append(relmse(native.finalize(sc.render_guided(cam, p, g, train=False)[0])[..., :3]))
        rp.append(relmse(native.finalize(sc.render(cam, p)[0])[..., :3]))
        ratios.append(rp[-1] / rg[-1])
    print("guided", [f"{x:.4g}" for x in rg], "plain", [f"{x:.4g}" for x in rp],
          "median gain", f"{np.median(ratios):.3f}")
    assert np.median(ratios) >= 0.8
    g.close()
    sc.close()


def test_guided_chunk_pairs_deterministic(ctx):
    """Several chunks: the chunks of a pair render concurrently with the
    same weights and the guide trains on both after the pair; the same
    seeds give the same image bit for bit and the same training steps,
    and the image agrees with plain rendering within 3 standard errors
    of the whole-image mean."""
    sd, cam = scenes.cornell_box()
    sc = native.Scene(ctx, sd)
    W, spp = 48, 32
    pif = (W + 4) * (W + 4) * 8          # 4 chunks of 8 spp
    outs = []
    for _ in range(2):
        g = native.Guide(ctx, seed=4)
        acc, _ = sc.render_guided(cam, abi.make_params(W, W, spp, max_path_len=8, seed=9,
                                                       paths_in_flight=pif), g)
        outs.append((acc, g.stats()["train_steps"]))
        g.close()
    (a, ta), (b, tb) = outs
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    assert ta == tb > 0
    K = 8
    gm = []
    for s in range(K):
        g = native.Guide(ctx, seed=60 + s)
        p = abi.make_params(W, W, spp, max_path_len=8, seed=50 + s, paths_in_flight=pif)
        gm.append(native.finalize(sc.render_guided(cam, p, g)[0])[..., :3].mean())
        g.close()
    pm = [native.finalize(sc.render(cam, abi.make_params(W, W, spp, max_path_len=8,
                                                          seed=80 + s))[0])[..., :3].mean()
          for s in range(K)]
    se = math.sqrt(np.var(gm, ddof=1) / K + np.var(pm, ddof=1) / K)
    assert abs(np.mean(gm) - np.mean(pm)) < 3 * se + 1e-6, (np.mean(gm), np.mean(pm), se)
    sc.close()
