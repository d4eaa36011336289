"""The photon-mapping integrator in the oracle (render/photon-integ.cc,
photon/*.cc), pinned without retyping its formulas:

- the radical inverse against exact digit reversal;
- the kd-tree: every node splits its subtree on the subtree's widest
  axis with the left side below and the right side above; the k-nearest
  search equals a brute force over all photons but the one a
  left-balanced tree of even size never searches (the reference's
  2i + 2 < n test), with the radius shrunk to the k-th distance;
- the direction histogram: bin pdfs integrate to 1, samples carry the
  pdf their bin has, and their frequencies follow the weights;
- the free light samples and the photon powers: the direct photons'
  power per path is each light's emitted flux (pi A Le for surface lights,
  4 pi^2 a^2 Le from a uniform sky onto a sphere of radius a);
- the density estimate: photons of power P on a jittered grid of
  spacing h over a Lambert floor give rho / pi * P / h^2 (up to the
  Gaussian filter's normalisation, 2.90898 against its exact 2.9110);
- end to end: a Cornell box with a glass ball, caustics and final
  gathering, renders to the path tracer's mean on the same camera
  samples within 2 %."""
import math
from fractions import Fraction

import numpy as np
import pytest

import oracle
from snogray_amd import abi, native, scenes
from tests.test_oracle_pins import IDENTITY_FRAME
from tests.test_oracle_shading import unit


def test_radical_inverse_closed_form():
    for base in (2, 3, 5, 7, 11, 13, 17):
        for n in (0, 1, 2, 3, 7, 10, 123, 4567, 99991, 1 << 20, 123456789):
            x, f, m = Fraction(0), Fraction(1, base), n
            while m:
                x += (m % base) * f
                m //= base
                f /= base
            assert abs(oracle.radical_inverse(n, base) - float(x)) < 1e-15, (n, base)


def _subtree(i, n):
    out, st = [], [i]
    while st:
        j = st.pop()
        if j < n:
            out.append(j)
            st += [2 * j + 1, 2 * j + 2]
    return out


@pytest.mark.parametrize("n", [1, 2, 7, 64, 301, 1000])
def test_kd_tree_splits(n):
    rng = np.random.default_rng(n)
    pts = rng.normal(size=(n, 3)) * np.array([3.0, 1.0, 0.5])
    kp, ax, _, _ = oracle.kd_probe(pts, np.zeros((1, 3)), 4, 1.0)
    # the same photons, in heap order
    assert sorted(map(tuple, kp.tolist())) == sorted(map(tuple, pts.astype(np.float32).tolist()))
    for i in range(n):
        if 2 * i + 1 >= n:
            continue
        sub = kp[_subtree(i, n)]
        ext = sub.max(0) - sub.min(0)
        widest = int(np.argmax(ext)) if ext.max() > 0 else 0
        assert ax[i] == widest
        a = ax[i]
        left, right = _subtree(2 * i + 1, n), _subtree(2 * i + 2, n)
        assert (kp[left, a] <= kp[i, a]).all()
        if right:
            assert (kp[right, a] >= kp[i, a]).all()


@pytest.mark.parametrize("n", [999, 1000])
def test_k_nearest_matches_brute_force(n):
    rng = np.random.default_rng(7 + n)
    pts = rng.uniform(-1, 1, (n, 3)).astype(np.float32)
    q = rng.uniform(-1.2, 1.2, (200, 3)).astype(np.float32)
    k, r2 = 12, 0.09
    kp, _, found, fr2 = oracle.kd_probe(pts, q, k, r2)
    # a heap of even size has one node (the last) under a parent with a
    # left child only: it is never searched
    reach = np.ones(n, bool)
    if n % 2 == 0:
        reach[n - 1] = False
    d2 = ((kp[None, :, :].astype(np.float64) - q[:, None, :]) ** 2).sum(2)
    for j in range(len(q)):
        cand = np.where(reach & (d2[j] < r2))[0]
        want = set(cand[np.argsort(d2[j][cand], kind="stable")][:k].tolist())
        got = set(found[j][found[j] >= 0].tolist())
        assert got == want, j
        if len(got) == k:
            assert abs(fr2[j] - max(d2[j][list(got)])) <= 1e-6 * fr2[j]
        else:
            assert fr2[j] == np.float32(r2)
    if n % 2 == 0:  # the lone left child is never found
        assert not (found == n - 1).any()


def test_direction_histogram_sampling():
    rng = np.random.default_rng(3)
    d = rng.normal(size=(400, 3)) + np.array([0.0, 0.0, 1.5])
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    w = rng.uniform(0.1, 2.0, 400)
    m = 40000
    par = rng.uniform(0, 1, (m, 2))
    out, bins = oracle.dirhist_probe(d, w, par)
    assert abs(bins.sum() / 64 - 1) < 1e-5                       # integrates to 1
    ok = out[:, 3] > 0
    assert ok.all()
    same = np.abs(out[:, 4] - out[:, 3]) <= 1e-6 * out[:, 3]
    assert same.mean() > 0.999                                   # pdf of the sampled bin
    # bin frequencies follow the weights: bin of each direction by its
    # sphere-sample parameters (u from the azimuth of (y, -x), v = (z + 1) / 2)
    def bin_of(v):
        phi = np.arctan2(v[:, 1], -v[:, 0])
        u = np.clip(0.5 - phi / (2 * np.pi), 0, 1)
        vv = np.clip((v[:, 2] + 1) / 2, 0, 1)
        return np.minimum((u * 8).astype(int), 7) + 8 * np.minimum((vv * 8).astype(int), 7)
    wb = np.bincount(bin_of(d), weights=w, minlength=64)
    sb = np.bincount(bin_of(out[:, :3].astype(float)), minlength=64)
    expect = wb / wb.sum() * m
    big = expect > 200
    assert np.all(np.abs(sb[big] - expect[big]) < 5 * np.sqrt(expect[big]))
    assert sb[wb == 0].sum() < 1e-3 * m


def _enclosed(light, size=20.0):
    """A black (albedo 0) Lambert box of side SIZE around LIGHT (a
    callable adding it to the description): every photon leaving the
    light is deposited once, directly."""
    sd = scenes.SceneDesc()
    blk = sd.lambert((0.0, 0.0, 0.0))
    h = size / 2
    for a in range(3):
        for s in (-1, 1):
            n = np.zeros(3)
            n[a] = s * h
            u = np.zeros(3)
            u[(a + 1) % 3] = size
            v = np.zeros(3)
            v[(a + 2) % 3] = size
            sd.tripar(tuple(n - u / 2 - v / 2), tuple(u), tuple(v), blk, parallelogram=True)
    light(sd)
    return sd


def _flux(sd, n=20000):
    pp = abi.photon_params(caustic=0, direct=n, indirect=0, direct_illum=False, seed=5)
    pm = oracle.OracleScene(sd).photon_map(pp)
    info = pm.info()
    pos, pw, _ = pm.export(0)
    return pw.astype(np.float64).sum(0) / info["paths"][0], info


LE = 3.0


@pytest.mark.parametrize("kind", ["tripar", "sphere", "ellipse", "cylinder", "sphere2"])
def test_photon_power_is_the_lights_flux(kind):
    def add(sd):
        g = sd.glow((LE, 2.0, 1.0))
        if kind == "tripar":
            sd.tripar((-0.5, 1.0, -0.4), (0, 0, 0.8), (1.2, 0, 0), g, parallelogram=True)
        elif kind == "sphere":
            sd.sphere((0.3, -0.2, 0.1), 0.7, g)
        elif kind == "ellipse":
            sd.ellipse((0.0, 0.5, 0.0), (0.6, 0.0, 0.0), (0.0, 0.0, 0.35), g)
        elif kind == "cylinder":
            sd.cylinder(native.cylinder_xform((0.0, -1.0, 0.0), (0.0, 0.0, 2.0), 0.4), g)
        else:
            sd.sphere2(np.hstack([0.55 * np.eye(3), np.array([[0.2], [0.1], [-0.3]])]), g)
    area = {"tripar": 0.8 * 1.2, "sphere": 4 * np.pi * 0.49, "ellipse": np.pi * 0.6 * 0.35,
            "cylinder": 2 * np.pi * 0.4 * 2.0, "sphere2": 4 * np.pi * 0.55 ** 2}[kind]
    flux, info = _flux(_enclosed(add))
    assert info["paths"][0] == info["photons"][0]     # every path leaves one photon
    np.testing.assert_allclose(flux, np.pi * area * np.array([LE, 2.0, 1.0]), rtol=2e-3)


def test_sky_photon_power_onto_a_sphere():
    """A uniform sky over a lone sphere of radius a: the sky's photons
    start on the disk tangent to the scene's bounding sphere, and those
    that hit carry 4 pi^2 a^2 Le per path in all."""
    a, L = 0.8, 2.5
    sd = scenes.SceneDesc()
    sd.sphere((0.1, 0.2, -0.1), a, sd.lambert((0.0, 0.0, 0.0)))
    sd.envmap(np.full((32, 64, 3), L, np.float32), frame=IDENTITY_FRAME)
    flux, info = _flux(sd, 40000)
    p = a * a / 3 / a ** 2  # hit fraction: pi a^2 over the disk of radius sqrt(3) a
    se = math.sqrt((1 - p) / (p * 40000))
    np.testing.assert_allclose(flux, 4 * np.pi ** 2 * a * a * L, rtol=4 * se)


def test_density_estimate_of_uniform_photons():
    """Photons of power P from above on a jittered grid of spacing h over
    a Lambert floor (albedo rho), as the direct map with direct lighting
    off and no other map: radiance rho / pi * P / h^2 times the filter's
    normalisation 2.90898 / 2.91104."""
    rho, P, h = 0.6, 0.02, 0.02
    rng = np.random.default_rng(11)
    g = np.arange(-1.5, 1.5, h)
    x, z = np.meshgrid(g, g, indexing="ij")
    pts = np.stack([x.ravel() + rng.uniform(0, h, x.size), np.zeros(x.size),
                    z.ravel() + rng.uniform(0, h, x.size)], 1).astype(np.float32)
    kp, ax, _, _ = oracle.kd_probe(pts, np.zeros((1, 3)), 1, 1.0)
    n = len(kp)
    pa = np.zeros((n, 4), np.float32)
    pa[:, :3] = kp
    pa[:, 3] = ax.astype(np.uint32).view(np.float32)
    pw = np.full((n, 3), P, np.float32)
    dr = np.tile(np.array([0.0, 1.0, 0.0], np.float32), (n, 1))
    empty = (np.zeros((0, 4), np.float32), np.zeros((0, 3), np.float32), np.zeros((0, 3), np.float32))
    pp = abi.photon_params(use_photons=50, radius=0.3, final_gather=False, direct_illum=False)
    pm = oracle.OraclePhotonMap.import_arrays(pp, [(pa, pw, dr), empty, empty], [1, 0, 0])
    sd = scenes.SceneDesc()
    sd.tripar((-50, 0, -50), (100, 0, 0), (0, 0, 100), sd.lambert((rho,) * 3), parallelogram=True)
    cam = scenes.camera((0.0, 1.0, 0.0), (0, -1, 0), (0, 0, 1), (1, 0, 0), 0.3, 0.3)
    p = abi.make_params(8, 8, 4, photon_map=pm)
    _, smp, _ = oracle.OracleScene(sd).render(cam, p, samples=True)
    expect = rho / np.pi * P / h ** 2 * (2.90898 / 2.911038)
    np.testing.assert_allclose(smp[:, 2].mean(), expect, rtol=0.01)
    # twice the power per photon, twice the radiance; a scale of 1/2 halves it
    pm2 = oracle.OraclePhotonMap.import_arrays(pp, [(pa, 2 * pw, dr), empty, empty], [2, 0, 0])
    _, smp2, _ = oracle.OracleScene(sd).render(cam, abi.make_params(8, 8, 4, photon_map=pm2),
                                               samples=True)
    np.testing.assert_allclose(smp2[:, 2], smp[:, 2], rtol=1e-5)


def caustic_cornell():
    sd, cam = scenes.cornell_box()
    sd.sphere((-0.1, 0.9, 0.2), 0.55, sd.glass(1.5))
    return sd, cam


def test_photon_render_matches_path_tracer():
    """Direct lighting, caustic photons (through the glass ball) and final
    gathering over the indirect map against the path tracer on the same
    camera samples (same seed and spp, so the light seen directly adds
    the same to both): the mean difference per sample is within 2 % of the
    image mean (measured: 1.0 / -0.1 / -0.05 %, the density estimate's
    bias); without final gathering (density estimates alone) within 8 %."""
    sd, cam = caustic_cornell()
    osc = oracle.OracleScene(sd)
    ref = osc.render(cam, abi.make_params(24, 24, 16, max_path_len=14, seed=9), samples=True)[1]
    rm = ref[:, 2:5].astype(float).mean(0)
    pm = osc.photon_map(abi.photon_params(caustic=3000, direct=30000, indirect=30000, seed=3))
    info = pm.info()
    assert info["photons"] == [30000, 3000, 30000]
    assert info["fg_bsdf_samples"] == 8 and info["fg_photon_samples"] == 8
    got = osc.render(cam, abi.make_params(24, 24, 16, photon_map=pm, seed=9), samples=True)[1]
    assert np.array_equal(got[:, :2], ref[:, :2])
    d = got[:, 2:5].astype(float) - ref[:, 2:5].astype(float)
    assert np.all(np.abs(d.mean(0)) < 0.02 * rm + 3 * d.std(0) / np.sqrt(len(d))), d.mean(0) / rm
    pm2 = osc.photon_map(abi.photon_params(caustic=3000, direct=30000, indirect=30000, seed=3,
                                           final_gather=False))
    assert pm2.info()["photons"][0] == 0  # direct lighting on, no gathering: no direct map
    got2 = osc.render(cam, abi.make_params(24, 24, 16, photon_map=pm2, seed=9), samples=True)[1]
    d2 = got2[:, 2:5].astype(float) - ref[:, 2:5].astype(float)
    assert np.all(np.abs(d2.mean(0)) < 0.08 * rm), d2.mean(0) / rm
