"""Serving loop on the GPU: requests arriving while others run are
admitted and overlap on the device; every result is bit-identical to the
same job rendered alone; scenes are evicted and rebuilt under a small
scene cap; sampled jobs pass the 3-standard-error pixel test against the
CPU oracle."""
import numpy as np
import pytest

import oracle
from snogray_amd import native, serve, workloads

pytestmark = pytest.mark.gpu

CAP = 1 << 16


def oracle_pixel_check(job, desc, acc, window=8, seeds=24):
    osc = oracle.OracleScene(desc)
    ww, hh = min(window, job.width), min(window, job.height)
    x0, y0 = (job.width - ww) // 2, (job.height - hh) // 2
    imgs = []
    for k in range(seeds):
        p = job.params(window=(x0, y0, ww, hh))
        p.seed = 1000 + 7919 * k + job.seed % 1000
        a, _, _ = osc.render(job.camera(), p)
        imgs.append(serve.finalize_numpy(a[y0:y0 + hh, x0:x0 + ww]))
    frac, ok, ref = serve.pixel_test(serve.finalize_numpy(acc[y0:y0 + hh, x0:x0 + ww]), imgs)
    return {"pass_fraction": frac, "cpu_pass_fraction": ref, "passed": bool(ok)}


def small_requests(n=24, rate=60.0, seed=3):
    rng = np.random.default_rng(seed)
    pool = [workloads.SceneSpec(n_spheres=int(rng.integers(1, 4)), level=int(rng.integers(2, 4)),
                                seed=int(rng.integers(1 << 20)), env_w=128, env_h=64)
            for _ in range(4)]
    t, reqs = 0.0, []
    for i in range(n):
        t += rng.exponential(1 / rate)
        job = workloads.JobSpec(pool[int(rng.integers(4))], int(rng.integers(12, 40)),
                                int(rng.integers(12, 40)), int(rng.integers(1, 9)),
                                seed=int(rng.integers(1 << 30)), max_path_len=6)
        reqs.append((t, job))
    return reqs


def test_serving_results_equal_alone_and_pass_pixel_test(ctx):
    reqs = small_requests()
    srv = serve.Server(ctx, budget=2 << 30, max_scenes=2, path_capacity=CAP, check_every=6,
                       check_fn=oracle_pixel_check)
    try:
        rep = srv.run(reqs, keep_results=True)
    finally:
        srv.close()
    assert rep["jobs"] == len(reqs)
    assert rep["p50_ms"] <= rep["p99_ms"]
    assert rep["max_active_jobs"] > 1               # requests overlapped
    assert rep["peak_budget_bytes"] <= rep["budget_bytes"]
    assert len(rep["checks"]) == len(reqs) // 6
    assert all(c["passed"] for c in rep["checks"]), rep["checks"]
    cache = {}
    for idx, (_, job) in enumerate(reqs):
        if job.scene not in cache:
            sd, _ = job.scene.build()
            cache[job.scene] = native.Scene(ctx, sd)
        alone, _ = cache[job.scene].render(job.camera(), job.params(paths_in_flight=CAP))
        got = rep["results"][idx]
        assert np.array_equal(got.view(np.uint32), alone.view(np.uint32)), idx
    for sc in cache.values():
        sc.close()
