"""The batched, served and learned workloads at their own sizes, and the
statistical image test on the icosphere soup at high spp.

- 64 uneven jobs (resolution 64-1024 per side, 1-256 spp, 10 k - 2 M
  triangles) through the autotuned engine: every job equals its alone
  render, and windows of some jobs pass the 3-standard-error test
  against the CPU oracle.
- Poisson arrivals of that job distribution at 64 jobs/s (a few seconds
  of them): every sampled job equals its alone render, and every 50th
  passes the pixel test.
- The learned proposal on the Cornell box at 1920x1080: finite,
  deterministic, the same sample positions as the plain render, and
  unbiased against it on crops over seeds.
- Cook-Torrance, glass and envmap light at 256 spp: per-pixel means over
  seeds within 3 standard errors of the oracle's."""
import time

import numpy as np
import pytest

import oracle
from snogray_amd import abi, batch, native, scenes, serve, workloads
from tests import helpers

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)


def window_check(desc, job, acc, window=12, seeds=16):
    """Calibrated 3-SE test (serve.pixel_test) of the central WINDOW of a
    job's accumulators against SEEDS independent oracle renders."""
    osc = oracle.OracleScene(desc)
    ww, hh = min(window, job.width), min(window, job.height)
    x0, y0 = (job.width - ww) // 2, (job.height - hh) // 2
    imgs = []
    for k in range(seeds):
        p = job.params(window=(x0, y0, ww, hh))
        p.seed = 5000 + 7919 * k + job.seed % 1000
        a, _, _ = osc.render(job.camera(), p)
        imgs.append(serve.finalize_numpy(a[y0:y0 + hh, x0:x0 + ww]))
    return serve.pixel_test(serve.finalize_numpy(acc[y0:y0 + hh, x0:x0 + ww]), imgs)


def test_batch_of_64_uneven_jobs(ctx):
    jobs = workloads.draw_jobs(64, 1)
    t0 = time.perf_counter()
    S, D, J = {}, {}, []
    for js in jobs:
        if js.scene not in S:
            D[js.scene] = js.scene.build()[0]
            S[js.scene] = native.Scene(ctx, D[js.scene])
        J.append((S[js.scene], js.camera(), js.params()))
    log(f"64 scenes built in {time.perf_counter() - t0:.1f} s, "
        f"{sum(j.samples for j in jobs) / 1e6:.0f} M samples")
    try:
        best, table = batch.autotune(ctx, J, log=log)
        log(f"tuned {best}")
        rep = batch.run_batch(ctx, J, keep_results=True, **best)
        log(f"batch: {rep['samples_per_s'] / 1e6:.0f} M samples/s, {rep['steps']} steps")
        assert rep["jobs"] == 64 and rep["max_jobs_per_step"] > 1
        same_chunks = best["path_capacity"] == abi.SGR_DEFAULT_PATHS_IN_FLIGHT
        for i, (sc, cam, p) in enumerate(J):
            alone, _ = sc.render(cam, p)
            got = rep["results"][i]
            if same_chunks:
                assert np.array_equal(got.view(np.uint32), alone.view(np.uint32)), i
            else:  # other chunking: the fp32 order of the chunk sums differs
                assert np.allclose(got, alone, rtol=2e-5, atol=1e-6), i
        log("every job equals its alone render")
        for i in (3, 17, 40):
            frac, ok, ref = window_check(D[jobs[i].scene], jobs[i], rep["results"][i])
            log(f"job {i} ({jobs[i].width}x{jobs[i].height}, {jobs[i].spp} spp, "
                f"{jobs[i].scene.triangles} triangles): pass {frac:.3f}, oracle {ref:.3f}")
            assert ok, (i, frac, ref)
    finally:
        for sc in S.values():
            sc.close()


def test_serving_poisson_64_jobs_per_s(ctx):
    duration, rate = 3.0, 64.0
    pool = workloads.scene_pool(12, 7)
    times = workloads.poisson_arrivals(rate, duration, 8)
    jobs = workloads.draw_jobs(len(times), 9, scene_pool=pool)
    descs = {}
    checks = []

    def check(job, desc, acc):
        frac, ok, ref = window_check(desc, job, acc)
        r = {"pass_fraction": frac, "cpu_pass_fraction": ref, "passed": bool(ok)}
        checks.append(r)
        return r

    srv = serve.Server(ctx, budget=64 << 30, max_scenes=8, check_every=50, check_fn=check)
    try:
        t_warm = srv.prewarm(pool)
        log(f"{len(jobs)} requests over {duration} s, prewarm {t_warm:.1f} s")
        rep = srv.run(list(zip(times, jobs)), keep_results=True, log=log)
        descs = dict(srv.host_descs)
    finally:
        srv.close()
    log(f"served: p50 {rep['p50_ms']:.0f} ms, p99 {rep['p99_ms']:.0f} ms, "
        f"max active {rep['max_active_jobs']}, checks {rep['checks']}")
    assert rep["jobs"] == len(jobs) and rep["unfinished"] == 0
    assert rep["max_active_jobs"] > 1
    assert rep["peak_budget_bytes"] <= rep["budget_bytes"]
    assert len(rep["checks"]) == len(jobs) // 50 and all(c["passed"] for c in rep["checks"])
    cache = {}
    try:
        for idx in range(0, len(jobs), 10):
            job = jobs[idx]
            if job.scene not in cache:
                cache[job.scene] = native.Scene(ctx, descs[job.scene])
            alone, _ = cache[job.scene].render(job.camera(), job.params())
            got = rep["results"][idx]
            assert np.array_equal(got.view(np.uint32), alone.view(np.uint32)), idx
        log(f"{len(range(0, len(jobs), 10))} sampled jobs equal their alone renders")
    finally:
        for sc in cache.values():
            sc.close()


def test_guided_cornell_1080p(ctx):
    W, H = 1920, 1080
    sd, cam = scenes.cornell_box(aspect=W / H)
    sc = native.Scene(ctx, sd)
    try:
        p = abi.make_params(W, H, 8, max_path_len=8, seed=21,
                            paths_in_flight=(W + 4) * (H + 4) * 2)
        imgs = []
        for _ in range(2):
            g = native.Guide(ctx, seed=5)
            acc, st = sc.render_guided(cam, p, g, train=True)
            g.close()
            imgs.append(acc)
        log(f"guided 1080p frame: {st['ms_render']:.1f} ms")
        assert np.isfinite(imgs[0]).all()
        assert np.array_equal(imgs[0].view(np.uint32), imgs[1].view(np.uint32))
        plain, _ = sc.render(cam, p)
        # the filter weights depend on the sample positions only, except
        # where the negative-lobe clamp rescales a tap by its radiance
        w_same = np.isclose(imgs[0][..., 4], plain[..., 4], rtol=1e-5, atol=1e-6)
        assert w_same.mean() > 0.99
        assert abs(imgs[0][..., 4].sum() / plain[..., 4].sum() - 1) < 1e-3
        # unbiased: crop means over seeds within 3 standard errors of plain
        win = (900, 500, 24, 16)
        gm, pm = [], []
        for s in range(12):
            q = abi.make_params(W, H, 16, max_path_len=8, seed=100 + s, window=win)
            g = native.Guide(ctx, seed=5)
            a, _ = sc.render_guided(cam, q, g, train=True)
            g.close()
            b, _ = sc.render(cam, abi.make_params(W, H, 16, max_path_len=8, seed=300 + s,
                                                  window=win))
            x0, y0, ww, hh = win
            gm.append(serve.finalize_numpy(a[y0:y0 + hh, x0:x0 + ww]).mean())
            pm.append(serve.finalize_numpy(b[y0:y0 + hh, x0:x0 + ww]).mean())
        gm, pm = np.array(gm), np.array(pm)
        se = np.sqrt(gm.var(ddof=1) / len(gm) + pm.var(ddof=1) / len(pm))
        log(f"crop mean guided {gm.mean():.4f} plain {pm.mean():.4f} se {se:.4f}")
        assert abs(gm.mean() - pm.mean()) < 3 * se
    finally:
        sc.close()


def test_soup_pixel_means_within_3se_of_oracle_at_256spp(ctx):
    """Cook-Torrance and glass spheres under the Gaussian-lobe envmap:
    per-pixel means of 12 GPU renders and 12 oracle renders (independent
    seeds, 256 spp each) agree within 3 combined standard errors."""
    sd, cam = helpers.small_soup(level=4, n_spheres=6, seed=3)
    sc, osc = native.Scene(ctx, sd), oracle.OracleScene(sd)
    W, H, spp = 48, 32, 256
    try:
        def g(seed):
            acc, _ = sc.render(cam, abi.make_params(W, H, spp, max_path_len=10, seed=seed))
            return serve.finalize_numpy(acc)

        def o(seed):
            acc, _, _ = osc.render(cam, abi.make_params(W, H, spp, max_path_len=10, seed=seed))
            return serve.finalize_numpy(acc)

        gi = np.stack([g(100 + s) for s in range(12)])
        t0 = time.perf_counter()
        oi = np.stack([o(900 + s) for s in range(12)])
        log(f"oracle renders {time.perf_counter() - t0:.1f} s")
        z = np.abs(gi.mean(0) - oi.mean(0)) / np.sqrt(gi.var(0, ddof=1) / 12 +
                                                      oi.var(0, ddof=1) / 12 + 1e-12)
        frac = float(np.mean(z <= 3))
        log(f"pixels within 3 SE: {frac:.4f}")
        assert frac > 0.97
    finally:
        sc.close()
