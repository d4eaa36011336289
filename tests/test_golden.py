"""Golden fixtures (tests/golden/, written by make_fixtures.py from the
oracle): the oracle must keep reproducing them exactly, the host octree
builder must give the same statistics, and the GPU must match the hits
bit for bit and the per-path samples to rounding."""
import json
import os

import numpy as np
import pytest

import oracle
from snogray_amd import abi, native, scenes
from tests import helpers

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def scene(name):
    if name == "cornell":
        return scenes.cornell_box()[0]
    return helpers.small_soup(level=4, n_spheres=3)[0]


def batches(z):
    k = 0
    while f"b{k}_rays" in z:
        r, s = z[f"b{k}_rays"], z[f"b{k}_shadow"]
        yield (k, abi.RayBatch(r[:3].T, r[3:6].T, r[6], r[7]),
               abi.RayBatch(s[:3].T, s[3:6].T, s[6], s[7]))
        k += 1


def assert_hits(h, z, k):
    assert np.array_equal(h.prim, z[f"b{k}_prim"])
    hit = h.prim != abi.MISS
    for f in ("t", "u", "v"):
        assert np.array_equal(getattr(h, f)[hit].view(np.uint32), z[f"b{k}_{f}"][hit].view(np.uint32))


@pytest.mark.parametrize("name", ["cornell", "soup_small"])
def test_oracle_reproduces_ray_fixtures(name):
    z = np.load(os.path.join(GOLD, f"rays_{name}.npz"))
    osc = oracle.OracleScene(scene(name))
    for k, b, sb in batches(z):
        h, _ = osc.trace_closest(b)
        assert_hits(h, z, k)
        occ, _ = osc.trace_occluded(sb)
        assert np.array_equal(occ, z[f"b{k}_occ"])


def test_oracle_reproduces_samples_and_bsdfs():
    z = np.load(os.path.join(GOLD, "samples_cornell_16x12x4.npz"))
    sd, cam = scenes.cornell_box()
    _, smp, _ = oracle.OracleScene(sd).render(cam, abi.make_params(16, 12, 4, max_path_len=8,
                                                                   seed=21), samples=True)
    assert np.array_equal(smp, z["samples"])
    b = np.load(os.path.join(GOLD, "bsdf_probe.npz"))
    mats = {
        "lambert": {"type": abi.SGR_MAT_LAMBERT, "color": (0.73, 0.5, 0.1)},
        "ct_glossy": {"type": abi.SGR_MAT_COOK_TORRANCE, "color": (0.2, 0.3, 0.1),
                      "gloss_color": (0.5, 0.5, 0.5), "m": 0.2, "ior_n": 1.6},
        "ct_rough": {"type": abi.SGR_MAT_COOK_TORRANCE, "color": (0.3, 0.3, 0.3),
                     "gloss_color": (0.5, 0.4, 0.3), "m": 0.6, "ior_n": 1.4},
        "glass": {"type": abi.SGR_MAT_GLASS, "ior_n": 1.5},
        "mirror": {"type": abi.SGR_MAT_MIRROR, "gloss_color": (0.9, 0.8, 0.7), "ior_n": 1.5},
    }
    for k, m in mats.items():
        s, e = oracle.bsdf_probe(m, b["v"], (0, 0, 1), b["params"], b["dirs"])
        assert np.array_equal(s, b[f"{k}_sample"]) and np.array_equal(e, b[f"{k}_eval"]), k


def test_host_octree_statistics_match_fixture():
    want = json.load(open(os.path.join(GOLD, "octree_stats.json")))
    descs = {"cornell": scenes.cornell_box()[0], "furnace": scenes.furnace_box()[0],
             "soup_small": helpers.small_soup(level=4, n_spheres=3)[0]}
    for name, sd in descs.items():
        _, _, _, st = native.octree_build(sd)
        for k, v in want[name].items():
            assert st[k] == pytest.approx(v, rel=1e-6), (name, k)


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["cornell", "soup_small"])
def test_gpu_matches_ray_fixtures(ctx, name):
    z = np.load(os.path.join(GOLD, f"rays_{name}.npz"))
    sc = native.Scene(ctx, scene(name))
    for k, b, sb in batches(z):
        h, _ = sc.trace_closest(b)
        assert_hits(h, z, k)
        occ, _ = sc.trace_occluded(sb)
        assert np.array_equal(occ, z[f"b{k}_occ"])
    sc.close()


@pytest.mark.gpu
def test_gpu_matches_sample_fixture(ctx):
    z = np.load(os.path.join(GOLD, "samples_cornell_16x12x4.npz"))
    sd, cam = scenes.cornell_box()
    sc = native.Scene(ctx, sd)
    got, _ = sc.render_samples(cam, abi.make_params(16, 12, 4, max_path_len=8, seed=21))
    ref = z["samples"]
    assert np.array_equal(got[:, :2], ref[:, :2])
    rel = np.abs(got[:, 2:] - ref[:, 2:]) / (np.abs(ref[:, 2:]) + 1e-3)
    assert np.mean(rel.max(1) <= 1e-4) > 0.99
    sc.close()
