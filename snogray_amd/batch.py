"""Batched uneven jobs on the multi-job engine, with an autotuner for the
engine's queue and launch sizes: the path capacity (paths in flight per
step, i.e. the length of the ray, shade and shadow queues), the launch
width of the camera-ray, refill-walk and splat kernels, and the
ray-queue length below which a late bounce's paths are finished in one
launch.

The tuner times a few engine steps of a probe batch (the real jobs with
their sample counts cut down) per candidate and keeps the configuration
with the most paths per device-millisecond.  A job's chunk is
min(its own chunk, capacity / pattern pixels) samples, so tuner choices
change a job's result only in the fp32 summation order of its
accumulators (the launch width and finishing threshold not at all)."""
import itertools
import time

import numpy as np

from . import abi, native

DEFAULT_CAPACITIES = (1 << 22, 1 << 23, 1 << 24, 1 << 25, 1 << 26)
# 0: 16 workgroups per CU.  The width reaches raygen, the refill walks and
# the splat; the other kernels are capped at what stays resident.
DEFAULT_GRIDS = (0, 4096 * 2)
DEFAULT_FINISH = (1 << 16, 1 << 18, 1 << 20)


def _probe(ctx, jobs, capacity, grid, max_steps, spp_cap, finish_max=0):
    """Paths per device-ms of the probe batch: a first pass captures the
    step graphs, the second pass is timed."""
    eng = native.Engine(ctx, path_capacity=capacity, grid_blocks=grid, use_graph=True,
                        finish_max=finish_max)
    rate = 0.0
    try:
        for _ in range(2):
            for sc, cam, p in jobs:
                q = abi.RenderParams.from_buffer_copy(p)
                q.spp = min(q.spp, spp_cap)
                eng.submit(sc, cam, q)
            paths, ms = 0, 0.0
            for _ in range(max_steps):
                si = eng.step()
                if si["jobs"] == 0:
                    break
                paths += si["paths"]
                ms += si["ms_device"]
            eng.release_all()
            rate = paths / ms if ms > 0 else 0.0
    finally:
        eng.release_all()
    return rate


def autotune(ctx, jobs, capacities=DEFAULT_CAPACITIES, grids=DEFAULT_GRIDS,
             finish=DEFAULT_FINISH, steps=6, spp_cap=16, probe_jobs=24, log=None):
    """JOBS: [(native.Scene, camera, params)].  Capacity and launch width
    are searched together, then the finishing threshold at the chosen
    pair.  Returns (best config dict, [(config, paths per ms)])."""
    def npix(p):
        _, _, w, h, m = abi.render_window(p)
        return (w + 2 * m) * (h + 2 * m)
    max_npix = max(npix(p) for _, _, p in jobs)
    probe = jobs[:probe_jobs]
    table = []
    for cap, grid in itertools.product(capacities, grids):
        if cap < max_npix:
            continue
        rate = _probe(ctx, probe, cap, grid, steps, spp_cap)
        table.append(({"path_capacity": cap, "grid_blocks": grid, "finish_max": 0}, rate))
        if log:
            log(f"autotune capacity {cap} grid {grid or 'auto'}: {rate / 1e3:.1f} M paths/s")
    best = pick(table)
    stage2 = [(c, r) for c, r in table if c == best]
    for fm in finish:
        cfg = dict(best, finish_max=fm)
        rate = _probe(ctx, probe, cfg["path_capacity"], cfg["grid_blocks"], steps, spp_cap, fm)
        table.append((cfg, rate))
        stage2.append((cfg, rate))
        if log:
            log(f"autotune finish below {fm} rays: {rate / 1e3:.1f} M paths/s")
    return max(stage2, key=lambda t: t[1])[0], table


def pick(table, tol=0.03):
    """The largest path capacity among the configurations within TOL of
    the best probe rate (the best rate among those).  A probe whose work
    fits the smaller capacities measures them alike, while the whole
    batch then needs more steps at the smaller one."""
    top = max(r for _, r in table)
    near = [(c, r) for c, r in table if r >= (1.0 - tol) * top]
    return max(near, key=lambda t: (t[0]["path_capacity"], t[1]))[0]


def run_batch(ctx, jobs, path_capacity=abi.SGR_DEFAULT_PATHS_IN_FLIGHT, grid_blocks=0,
              keep_results=False, finish_max=0):
    """Submit every job, step until all are done.  Returns a report with
    aggregate samples/s and per-job latency (submit to done)."""
    eng = native.Engine(ctx, path_capacity=path_capacity, grid_blocks=grid_blocks,
                        use_graph=True, finish_max=finish_max)
    results = {}
    t0 = time.perf_counter()
    ids = [eng.submit(sc, cam, p) for sc, cam, p in jobs]
    steps = eng.run()
    wall = time.perf_counter() - t0
    lat, samples = [], 0
    for i, (jid, (_, _, p)) in enumerate(zip(ids, jobs)):
        st = eng.poll(jid)
        lat.append(st["t_done_ms"] - st["t_submit_ms"])
        x0, y0, w, h, _ = abi.render_window(p)
        samples += w * h * p.spp
        if keep_results:
            results[i] = eng.result(jid, p.width, p.height)
    stats = eng.stats()
    eng.release_all()
    lat = np.array(lat)
    rep = {"jobs": len(jobs), "wall_s": wall, "samples": int(samples),
           "samples_per_s": samples / wall,
           "device_ms": float(sum(s["ms_device"] for s in steps)),
           "steps": len(steps), "graph_steps": int(sum(s["graph"] for s in steps)),
           "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
           "max_jobs_per_step": int(max(s["jobs"] for s in steps)),
           "scene_intersect_calls": stats["scene_intersect_calls"],
           "shadow_tests": stats["shadow_tests"]}
    if keep_results:
        rep["results"] = results
    return rep
