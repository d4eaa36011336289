"""sgr_render_noalloc: the render entry point that allocates no device
memory (its path state lives in a fixed per-device scratch array) gives
the image sgr_render gives with the same chunking, on two calls in a row
(the second smaller, on the scratch the first left behind), and rejects
a window too large for its scratch."""
import ctypes as C

import numpy as np
import pytest

from snogray_amd import abi, native, scenes
from tests import helpers

pytestmark = pytest.mark.gpu


def noalloc(sc, cam, p):
    dev = helpers.DeviceMem(p.width * p.height * 5 * 4)
    try:
        st = abi.Stats()
        rc = native.lib().sgr_render_noalloc(sc.h, C.byref(cam), C.byref(p),
                                             C.cast(C.c_void_p(dev.ptr), abi.f32p),
                                             C.byref(st))
        if rc:
            raise native.SgrError(native.lib().sgr_last_error().decode())
        return dev.to_host((p.height, p.width, 5)), st.as_dict()
    finally:
        dev.free()


def pattern_pixels(p):
    _, _, w, h, m = abi.render_window(p)
    return (w + 2 * m) * (h + 2 * m)


def test_noalloc_matches_render_twice(ctx):
    sd, cam = scenes.cornell_box()
    sc = native.Scene(ctx, sd)
    try:
        for (w, h, spp, per) in ((64, 48, 16, 4), (24, 20, 8, 2)):
            p = abi.make_params(w, h, spp, max_path_len=8)
            p.paths_in_flight = pattern_pixels(p) * per
            print(f"noalloc {w}x{h} spp {spp}, {per} samples per chunk", flush=True)
            got, gst = noalloc(sc, cam, p)
            ref, rst = sc.render(cam, p)
            assert np.array_equal(got, ref)
            for k in ("eye_rays", "scene_intersect_calls", "shadow_tests", "node_visits_closest"):
                assert gst[k] == rst[k], k
            assert got[..., 4].min() > 0
    finally:
        sc.close()


def test_noalloc_chunks_to_fit_and_rejects_too_large(ctx):
    sd, cam = scenes.cornell_box()
    sc = native.Scene(ctx, sd)
    try:
        # default paths in flight: the scratch sets the chunk; only the fp32
        # order of the per-chunk partial sums can differ
        p = abi.make_params(96, 80, 32, max_path_len=8)
        print("noalloc 96x80 spp 32, default chunking", flush=True)
        got, _ = noalloc(sc, cam, p)
        ref, _ = sc.render(cam, p)
        np.testing.assert_allclose(got, ref, rtol=2e-5, atol=1e-5)
        big = abi.make_params(1024, 1024, 1, max_path_len=8)
        with pytest.raises(native.SgrError, match="scratch holds"):
            noalloc(sc, cam, big)
    finally:
        sc.close()
