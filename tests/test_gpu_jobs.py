"""Multi-job engine on the GPU: a job's accumulators must be bit-identical
to sgr_render's with the same parameters whether it runs alone, in a
mixed batch, from captured step graphs or eager launches; admission must
respect the memory budget; segments beyond one step's table roll over."""
import numpy as np
import pytest

from snogray_amd import abi, native, scenes
from tests import helpers

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def job_mix(ctx):
    """Three scenes, six jobs with different sizes, windows, sample counts,
    path lengths, light sampling and chunking."""
    cb, ccam = scenes.cornell_box()
    fb, fcam = scenes.furnace_box()
    ss, scam = helpers.small_soup(level=4, n_spheres=4, seed=11)
    S = [native.Scene(ctx, d) for d in (cb, fb, ss)]
    cams = [ccam, fcam, scam]
    spec = [
        (0, abi.make_params(40, 30, 6, max_path_len=6, seed=3, paths_in_flight=2 * 34 * 44)),
        (1, abi.make_params(24, 24, 3, max_path_len=4, seed=5)),
        (2, abi.make_params(64, 36, 5, max_path_len=9, seed=7, paths_in_flight=3 * 68 * 40)),
        (0, abi.make_params(40, 30, 4, max_path_len=8, seed=9, direct_samples=2,
                            window=(8, 6, 20, 12))),
        (2, abi.make_params(32, 20, 7, max_path_len=3, min_path_len=1, seed=13,
                            paths_in_flight=36 * 24)),
        (1, abi.make_params(16, 12, 2, max_path_len=10, seed=17)),
    ]
    jobs = [(S[i], cams[i], p) for i, p in spec]
    alone = [sc.render(cam, p)[0] for sc, cam, p in jobs]
    yield jobs, alone
    for s in S:
        s.close()


def run_engine(ctx, jobs, **cfg):
    eng = native.Engine(ctx, **cfg)
    ids = [eng.submit(sc, cam, p) for sc, cam, p in jobs]
    steps = eng.run()
    outs = []
    for jid, (_, _, p) in zip(ids, jobs):
        st = eng.poll(jid)
        assert st["state"] == abi.SGR_JOB_DONE
        assert st["chunks_done"] == st["chunks_total"]
        assert 0 <= st["t_submit_ms"] <= st["t_start_ms"] <= st["t_done_ms"]
        outs.append(eng.result(jid, p.width, p.height))
        eng.release(jid)
    return outs, steps


def assert_bits(a, b):
    assert a.shape == b.shape
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_batch_equals_alone_graph(ctx, job_mix):
    jobs, alone = job_mix
    # capacity for about two jobs' chunks: steps mix jobs and rotate
    outs, steps = run_engine(ctx, jobs, path_capacity=12000, use_graph=True)
    assert any(s["jobs"] > 1 for s in steps)
    assert all(s["graph"] for s in steps)
    assert all(s["paths"] <= 12000 for s in steps)
    for o, a in zip(outs, alone):
        assert_bits(o, a)


def test_graph_replay_equals_eager(ctx, job_mix):
    jobs, alone = job_mix
    eager, st_e = run_engine(ctx, jobs, path_capacity=1 << 20, use_graph=False)
    assert not any(s["graph"] for s in st_e)
    # several replays of each captured graph, eager work between them
    for _ in range(2):
        graphed, _ = run_engine(ctx, jobs, path_capacity=1 << 20, use_graph=True)
        for g, e, a in zip(graphed, eager, alone):
            assert_bits(g, e)
            assert_bits(g, a)
        sc, cam, p = jobs[1]
        assert_bits(sc.render(cam, p)[0], alone[1])


def test_render_batch(ctx, job_mix):
    jobs, alone = job_mix
    native.Engine(ctx)  # default configuration
    outs, st = native.render_batch(ctx, jobs)
    for o, a in zip(outs, alone):
        assert_bits(o, a)
    assert st["samples"] == sum(p.spp * abi.render_window(p)[2] * abi.render_window(p)[3]
                                for _, _, p in jobs)


def test_admission_respects_budget(ctx, job_mix):
    jobs, alone = job_mix
    fps = [native.Engine.footprint(p) for _, _, p in jobs]
    budget = max(fps) + min(fps)
    eng = native.Engine(ctx, memory_budget=budget)
    ids = [eng.submit(sc, cam, p) for sc, cam, p in jobs]
    done = {}
    while len(done) < len(ids):
        si = eng.step()
        held = 0
        for jid, fp in zip(ids, fps):
            if jid in done:
                continue
            st = eng.poll(jid)
            if st["state"] != abi.SGR_JOB_QUEUED:
                held += st["footprint"]
            if st["state"] == abi.SGR_JOB_DONE:
                _, _, p = jobs[ids.index(jid)]
                done[jid] = eng.result(jid, p.width, p.height)
                eng.release(jid)
        assert held <= budget
        assert si["jobs"] > 0 or done
    for jid, a in zip(ids, alone):
        assert_bits(done[jid], a)


def test_many_tiny_jobs_roll_over_segments(ctx):
    """More jobs than one step's segment table (256)."""
    sd, cam = scenes.cornell_box()
    sc = native.Scene(ctx, sd)
    jobs = [(sc, cam, abi.make_params(6, 5, 1 + i % 3, max_path_len=4, seed=100 + i))
            for i in range(300)]
    outs, steps = run_engine(ctx, jobs)
    assert max(s["jobs"] for s in steps) == 256
    for i in (0, 1, 2, 255, 256, 299):
        s, c, p = jobs[i]
        assert_bits(outs[i], s.render(c, p)[0])
    sc.close()


def test_scene_in_use_cannot_be_destroyed(ctx):
    sd, cam = scenes.cornell_box()
    sc = native.Scene(ctx, sd)
    eng = native.Engine(ctx)
    jid = eng.submit(sc, cam, abi.make_params(8, 8, 1))
    assert native.lib().sgr_scene_destroy(sc.h) != 0
    assert b"job" in native.lib().sgr_last_error()
    with pytest.raises(native.SgrError):
        eng.result(jid, 8, 8)  # not finished
    eng.run()
    eng.release(jid)
    sc.close()


def test_engine_rejects_bad_jobs(ctx):
    sd, cam = scenes.cornell_box()
    sc = native.Scene(ctx, sd)
    eng = native.Engine(ctx, path_capacity=1000)
    with pytest.raises(native.SgrError):
        eng.submit(sc, cam, abi.make_params(64, 64, 1))  # 68*68 pixels > capacity
    with pytest.raises(native.SgrError):
        eng.submit(sc, cam, abi.make_params(0, 8, 1))
    eng2 = native.Engine(ctx, memory_budget=100)
    with pytest.raises(native.SgrError):
        eng2.submit(sc, cam, abi.make_params(8, 8, 1))  # 1280-byte accumulators
    with pytest.raises(native.SgrError):
        eng2.poll(12345)
    native.Engine(ctx)
    sc.close()
