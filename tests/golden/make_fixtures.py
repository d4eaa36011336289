"""Writes the golden fixtures of this directory from the CPU oracle.

    python tests/golden/make_fixtures.py

The reference cannot be built in this image (its sources include an
autoconf-generated config.h), so the fixtures are the oracle's outputs on
fixed, seeded inputs; the oracle itself is pinned by closed forms and
brute force (tests/test_oracle_*.py).  The fixtures freeze those outputs:
a change to the oracle or to the GPU path that alters a hit, a sample or
a BSDF value shows up against them.
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from snogray_amd import abi, scenes  # noqa: E402
from tests import helpers  # noqa: E402


def ray_fixture(name, sd, cam, n_random, seed):
    osc = oracle.OracleScene(sd)
    lo, hi = sd.bbox()
    lo, hi = np.asarray(lo), np.asarray(hi)
    pad = 0.2 * (hi - lo)
    batches = [helpers.random_rays(lo - pad, hi + pad, n_random, seed=seed)]
    if cam is not None:
        batches.append(helpers.camera_rays(cam, 48, 32, seed=seed + 1))
    out = {}
    for k, b in enumerate(batches):
        h, _ = osc.trace_closest(b)
        sb = helpers.shadow_batch(b, h)
        occ, _ = osc.trace_occluded(sb)
        out[f"b{k}_rays"] = np.stack(b.cols)       # ox oy oz dx dy dz t0 t1
        out[f"b{k}_shadow"] = np.stack(sb.cols)
        out[f"b{k}_prim"] = h.prim
        out[f"b{k}_t"] = h.t
        out[f"b{k}_u"] = h.u
        out[f"b{k}_v"] = h.v
        out[f"b{k}_occ"] = occ
    np.savez_compressed(os.path.join(HERE, f"rays_{name}.npz"), **out)


def samples_fixture():
    sd, cam = scenes.cornell_box()
    osc = oracle.OracleScene(sd)
    p = abi.make_params(16, 12, 4, max_path_len=8, seed=21)
    acc, smp, st = osc.render(cam, p, samples=True)
    np.savez_compressed(os.path.join(HERE, "samples_cornell_16x12x4.npz"), samples=smp,
                        accum=acc.astype(np.float32))


def bsdf_fixture():
    rng = np.random.default_rng(31)
    prm = rng.random((64, 2))
    dirs = rng.normal(size=(64, 3))
    dirs[:, 2] = np.abs(dirs[:, 2])
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    v = np.array([0.3, -0.2, 0.9])
    v /= np.linalg.norm(v)
    mats = {
        "lambert": {"type": abi.SGR_MAT_LAMBERT, "color": (0.73, 0.5, 0.1)},
        "ct_glossy": {"type": abi.SGR_MAT_COOK_TORRANCE, "color": (0.2, 0.3, 0.1),
                      "gloss_color": (0.5, 0.5, 0.5), "m": 0.2, "ior_n": 1.6},
        "ct_rough": {"type": abi.SGR_MAT_COOK_TORRANCE, "color": (0.3, 0.3, 0.3),
                     "gloss_color": (0.5, 0.4, 0.3), "m": 0.6, "ior_n": 1.4},
        "glass": {"type": abi.SGR_MAT_GLASS, "ior_n": 1.5},
        "mirror": {"type": abi.SGR_MAT_MIRROR, "gloss_color": (0.9, 0.8, 0.7), "ior_n": 1.5},
    }
    out = {"params": prm, "dirs": dirs, "v": v}
    for k, m in mats.items():
        s, e = oracle.bsdf_probe(m, v, (0, 0, 1), prm, dirs)
        out[f"{k}_sample"] = s
        out[f"{k}_eval"] = e
    np.savez_compressed(os.path.join(HERE, "bsdf_probe.npz"), **out)


def octree_fixture():
    stats = {}
    for name, (sd, _) in {"cornell": scenes.cornell_box(), "furnace": scenes.furnace_box(),
                          "soup_small": helpers.small_soup(level=4, n_spheres=3)}.items():
        st = oracle.OracleScene(sd).octree_stats()
        stats[name] = {k: (float(v) if isinstance(v, float) else int(v)) for k, v in st.items()}
    with open(os.path.join(HERE, "octree_stats.json"), "w") as f:
        json.dump(stats, f, indent=1, sort_keys=True)


def main():
    oracle.build()
    sd, cam = scenes.cornell_box()
    ray_fixture("cornell", sd, cam, 3000, 41)
    sd, cam = helpers.small_soup(level=4, n_spheres=3)
    ray_fixture("soup_small", sd, cam, 3000, 43)
    samples_fixture()
    bsdf_fixture()
    octree_fixture()


if __name__ == "__main__":
    main()
