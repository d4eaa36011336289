"""Host side of the product: the C-ABI library loads and exports every
symbol include/snoggpu.h declares, and the host octree builder
(snogray_amd/csrc/octree_build.cpp) produces exactly the oracle's
flattened tree -- node for node and reference for reference."""
import os
import re

import numpy as np
import pytest

import oracle
from snogray_amd import abi, native, scenes
from tests import helpers

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "snoggpu.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char \*)\s*(sgr_\w+)\(", text, re.M)))


def test_library_exports_every_declared_symbol():
    L = native.lib()
    syms = declared_symbols()
    assert len(syms) >= 16
    for s in syms:
        assert hasattr(L, s), s
    assert set(native.EXPORTS) <= set(syms)


def test_errors_are_reported_not_raised():
    """Bad descriptors fail with SGR_ERR_INVALID and a message
    (the reference throws std::runtime_error, render/global-render-state.cc:73)."""
    sd = scenes.SceneDesc()
    m = sd.lambert((0.5, 0.5, 0.5))
    sd.mesh(np.zeros((3, 3), np.float32), np.array([[0, 1, 5]], np.uint32), m)
    with pytest.raises(native.SgrError, match="vertex index out of range"):
        native.octree_build(sd)
    sd = scenes.SceneDesc()
    sd.mesh(np.zeros((3, 3), np.float32), np.array([[0, 1, 2]], np.uint32), 3)
    with pytest.raises(native.SgrError, match="bad material"):
        native.octree_build(sd)
    # a non-finite corner would grow the root without end
    for bad in (np.nan, np.inf):
        sd = scenes.SceneDesc()
        m = sd.lambert((0.5, 0.5, 0.5))
        v = np.array([(0, 0, 0), (1, 0, 0), (0, bad, 0)], np.float32)
        sd.mesh(v, np.array([[0, 1, 2]], np.uint32), m)
        with pytest.raises(native.SgrError, match="non-finite"):
            native.octree_build(sd)
    sd = scenes.SceneDesc()
    m = sd.lambert((0.5, 0.5, 0.5))
    sd.sphere((0, 0, 0), np.inf, m)
    with pytest.raises(native.SgrError, match="non-finite"):
        native.octree_build(sd)
    sd = scenes.SceneDesc()
    m = sd.lambert((0.5, 0.5, 0.5))
    sd.tripar((0, 0, 0), (1, 0, 0), (0, np.nan, 0), m)
    with pytest.raises(native.SgrError, match="non-finite"):
        native.octree_build(sd)


def assert_same_tree(sd):
    n1, s1, o1, st1 = native.octree_build(sd)
    o = oracle.OracleScene(sd)
    n2, s2, o2 = o.octree()
    assert np.array_equal(n1, n2)
    assert np.array_equal(s1, s2)
    assert np.array_equal(o1.astype(np.float64), o2)
    st2 = o.octree_stats()
    for k in ("num_nodes", "num_leaf_nodes", "num_surfaces", "num_dup_surfaces", "max_depth"):
        assert st1[k] == st2[k], k
    assert np.float32(st1["avg_depth"]) == np.float32(st2["avg_depth"])


@pytest.mark.parametrize("name", ["cornell", "furnace", "soup_rand", "icosoup_small",
                                  "icosoup_327k", "degenerate", "coincident"])
def test_host_octree_matches_oracle(name):
    if name == "cornell":
        sd, _ = scenes.cornell_box()
    elif name == "furnace":
        sd, _ = scenes.furnace_box()
    elif name == "soup_rand":
        sd = helpers.triangle_soup(5000, seed=21, extent=20.0, size=0.3)
    elif name == "icosoup_small":
        sd, _ = helpers.small_soup(level=4, n_spheres=4)
    elif name == "icosoup_327k":
        # large enough that the builder descends in parallel runs
        sd, _ = scenes.icosphere_soup(level=5)
    elif name == "degenerate":
        # zero-area and repeated triangles are skipped / kept like the
        # reference (surface/mesh.cc:832-846)
        sd = scenes.SceneDesc()
        m = sd.lambert((0.5, 0.5, 0.5))
        v = np.array([(0, 0, 0), (1, 0, 0), (0, 1, 0), (2, 2, 2), (3, 3, 3)], np.float32)
        sd.mesh(v, np.array([[0, 1, 2], [0, 0, 0], [3, 4, 4], [0, 1, 2]], np.uint32), m)
    else:
        # many surfaces sharing exact midpoints exercise the tie rules
        sd = scenes.SceneDesc()
        m = sd.lambert((0.5, 0.5, 0.5))
        g = np.arange(0, 9, dtype=np.float32)
        vs, ts = [], []
        for i in g:
            for j in g:
                base = len(vs)
                vs += [(i, j, 0), (i + 1, j, 0), (i, j + 1, 0)]
                ts.append([base, base + 1, base + 2])
        sd.mesh(np.array(vs, np.float32), np.array(ts, np.uint32), m)
        sd.tripar((0, 0, 0), (8, 0, 0), (0, 8, 0), m, parallelogram=True)
        sd.sphere((4, 4, 4), 2.0, m)
    assert_same_tree(sd)


@pytest.mark.slow
def test_host_octree_matches_oracle_full_soup():
    sd, _ = scenes.icosphere_soup()
    assert_same_tree(sd)
