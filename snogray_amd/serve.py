"""Serving loop: continuous batching of render requests on one GPU.

Requests arrive over time (a Poisson process in the benchmark).  A
request is admitted as soon as its scene is resident and its accumulators
fit the HBM budget; admitted requests join the job engine, whose every
step carries the next sample chunk of every active job, so requests that
arrive while others run overlap with them on the device.  Up to
``max_scenes`` scenes stay resident (least recently used idle scenes are
evicted when a new one needs the room); scenes are built on a host thread
while the GPU keeps stepping.  Every ``check_every``-th finished job is
handed, with its accumulators, to an optional check function (the
benchmark compares a window of it with the CPU renderer by the
3-standard-error pixel test) on a host thread.

``ResidencyPlanner`` is the admission bookkeeping alone (no device
calls), so its invariant -- resident scene bytes + admitted accumulator
bytes + reserved bytes never exceed the budget -- is testable anywhere.
"""
import collections
import concurrent.futures as cf
import time

import numpy as np

from . import abi, native


class ResidencyPlanner:
    """HBM-footprint admission: scenes (resident, LRU, at most
    MAX_SCENES) and job accumulators against BUDGET bytes, of which
    RESERVED (engine path buffers) are held from the start."""

    def __init__(self, budget, max_scenes=32, reserved=0):
        if reserved > budget:
            raise ValueError("reserved bytes exceed the budget")
        self.budget = int(budget)
        self.max_scenes = int(max_scenes)
        self.reserved = int(reserved)
        self.scenes = collections.OrderedDict()  # key -> [bytes, active jobs]
        self.job_bytes = 0
        self.peak = self.used()

    def used(self):
        return self.reserved + self.job_bytes + sum(v[0] for v in self.scenes.values())

    def resident(self, key):
        return key in self.scenes

    def _evictable(self, protect=()):
        return [k for k, v in self.scenes.items() if v[1] == 0 and k not in protect]

    def plan_scene(self, key, nbytes, protect=()):
        """Scenes to evict so that scene KEY of NBYTES fits (None if it
        cannot fit now).  Idle scenes are evicted oldest first."""
        if key in self.scenes:
            return []
        free = self.budget - self.used()
        slots = self.max_scenes - len(self.scenes)
        evict = []
        for k in self._evictable(protect):
            if free >= nbytes and slots >= 1:
                break
            evict.append(k)
            free += self.scenes[k][0]
            slots += 1
        if free >= nbytes and slots >= 1:
            return evict
        return None

    def make_room(self, nbytes, protect=()):
        """Idle scenes to evict so that NBYTES more fit (None if they
        cannot)."""
        free = self.budget - self.used()
        evict = []
        for k in self._evictable(protect):
            if free >= nbytes:
                break
            evict.append(k)
            free += self.scenes[k][0]
        return evict if free >= nbytes else None

    def add_scene(self, key, nbytes):
        assert key not in self.scenes
        assert len(self.scenes) < self.max_scenes
        assert self.used() + nbytes <= self.budget
        self.scenes[key] = [int(nbytes), 0]
        self._note()

    def resize_scene(self, key, nbytes):
        """Replace a scene's estimated size by its measured one; returns
        False (and changes nothing) if that would break the budget."""
        old = self.scenes[key][0]
        if self.used() - old + nbytes > self.budget:
            return False
        self.scenes[key][0] = int(nbytes)
        self._note()
        return True

    def evict(self, key):
        assert self.scenes[key][1] == 0, "scene still has active jobs"
        del self.scenes[key]

    def drain_target(self, n_starving):
        """How many scenes to keep draining while N_STARVING scenes that
        waiting jobs need find no room: one per starving scene, at most a
        quarter of the resident scenes (at least one), so that most
        scenes keep taking jobs while room is made."""
        return min(int(n_starving), max(1, len(self.scenes) // 4))

    def drain_victims(self, n, demand, draining=(), protect=()):
        """N resident scenes to stop admitting to, so that their active
        jobs finish and the scenes become evictable for scenes that
        waiting jobs need.  Scenes that no waiting job wants (DEMAND:
        key -> number of waiting jobs) go first, then the least
        recently admitted; scenes already DRAINING or in PROTECT are
        never chosen."""
        cand = [k for k in self.scenes if k not in draining and k not in protect]
        # stable sort: LRU order within each demand class
        cand.sort(key=lambda k: demand.get(k, 0) > 0)
        return cand[:max(0, int(n))]

    def can_admit(self, key, nbytes):
        return key in self.scenes and self.used() + nbytes <= self.budget

    def admit(self, key, nbytes):
        assert self.can_admit(key, nbytes)
        self.scenes[key][1] += 1
        self.scenes.move_to_end(key)
        self.job_bytes += int(nbytes)
        self._note()

    def release(self, key, nbytes):
        self.scenes[key][1] -= 1
        assert self.scenes[key][1] >= 0
        self.job_bytes -= int(nbytes)

    def _note(self):
        u = self.used()
        assert u <= self.budget, "budget exceeded"
        self.peak = max(self.peak, u)


def pixel_test(gpu_img, cpu_imgs, slack=0.03):
    """3-standard-error test per pixel and channel of GPU_IMG (one render
    at the job's spp) against CPU_IMGS (k renders at the same spp and
    independent seeds): |gpu - mean| <= 3 sqrt(sd^2 + sd^2 / k), sd the
    spread of one CPU render.

    At a few samples per pixel a render's pixels are heavy-tailed (a rare
    bright path), so the pass fraction of a correct renderer is below the
    Gaussian 99.7 %.  The test is therefore calibrated by the CPU renders
    themselves: each is tested the same way against the other k - 1 (leave
    one out), and the GPU render passes when its fraction is within SLACK
    (about 4 binomial standard errors of a 16 x 16 window) of their median.
    Returns (GPU pass fraction, passed, median CPU pass fraction)."""
    c = np.stack(cpu_imgs).astype(np.float64)
    k = len(c)

    def frac(x, others):
        mean = others.mean(0)
        sd = others.std(0, ddof=1)
        n = len(others)
        tol = 3 * np.sqrt(sd ** 2 + sd ** 2 / n) + 1e-6 * np.abs(mean) + 1e-7
        return float(np.mean(np.abs(x - mean) <= tol))

    g = frac(np.asarray(gpu_img, np.float64), c)
    ref = float(np.median([frac(c[i], np.delete(c, i, axis=0)) for i in range(k)]))
    return g, g >= ref - slack, ref


def finalize_numpy(acc):
    """ImageSampledOutput::set_raw_min_y normalisation of accumulators."""
    acc = np.asarray(acc, np.float64)
    w = acc[..., 4:5]
    col = np.where(w > 0, acc[..., :3] / np.where(w > 0, w, 1), acc[..., :3])
    a = np.clip(np.where(w[..., 0] > 0, acc[..., 3] / np.where(w[..., 0] > 0, w[..., 0], 1),
                         acc[..., 3]), 0, 1)
    col = np.maximum(col, 0)
    col = np.where((a > 0)[..., None] & (a < 1)[..., None], col / np.where(a > 0, a, 1)[..., None],
                   col)
    return col


class Server:
    """Drives one context's job engine for a stream of (arrival time,
    JobSpec) requests; see the module docstring."""

    def __init__(self, ctx, budget, max_scenes=32, path_capacity=1 << 26, check_every=50,
                 check_fn=None, use_graph=True, builders=6):
        self.ctx = ctx
        self.engine = native.Engine(ctx, path_capacity=path_capacity, use_graph=use_graph)
        # the engine's path buffers are held for the whole run
        per_path = 340 + 45 * 2
        self.planner = ResidencyPlanner(budget, max_scenes, reserved=per_path * path_capacity)
        self.scenes = {}        # SceneSpec -> native.Scene
        self.building = {}      # SceneSpec -> Future
        self.builder = cf.ThreadPoolExecutor(builders)
        self.host_descs = {}    # SceneSpec -> SceneDesc (host copies outlive eviction)
        self.checker = cf.ThreadPoolExecutor(1)
        self.check_every = check_every if check_fn else 0
        self.check_fn = check_fn  # (job, scene desc, accumulators) -> dict
        self.descs = {}
        # resident scenes that take no new jobs until their active jobs
        # finish and they can be evicted for a scene a waiting job needs
        self.draining = set()

    def _build(self, spec):
        sd = self.host_descs.get(spec)
        if sd is None:
            sd, _ = spec.build()
            self.host_descs[spec] = sd
        sc = native.Scene(self.ctx, sd)
        return sd, sc

    def prewarm(self, specs):
        """Make SPECS resident (up to the scene cap and budget) before
        requests start; returns the seconds it took."""
        t0 = time.perf_counter()
        futs = {}
        for spec in specs:
            if len(self.planner.scenes) >= self.planner.max_scenes:
                break
            if spec in self.planner.scenes:
                continue
            if self.planner.plan_scene(spec, spec.estimated_bytes()) != []:
                break
            self.planner.add_scene(spec, spec.estimated_bytes())
            futs[spec] = self.builder.submit(self._build, spec)
        for spec, f in futs.items():
            sd, sc = f.result()
            self.descs[spec] = sd
            self.scenes[spec] = sc
            if not self.planner.resize_scene(spec, sc.footprint()):
                raise RuntimeError("prewarmed scenes exceed the budget")
        return time.perf_counter() - t0

    def run(self, requests, drain=True, log=None, keep_results=False, drain_limit=None):
        """REQUESTS: list of (arrival seconds, JobSpec), sorted by time.
        Returns a report dict (with every job's accumulators under
        "results" if KEEP_RESULTS).  With DRAIN the loop runs until every
        request has finished, or at most DRAIN_LIMIT seconds after the
        last arrival; requests still unfinished then are counted under
        "unfinished"."""
        results = {}
        reqs = collections.deque((a, i, job) for i, (a, job) in enumerate(requests))
        pending = collections.deque()   # (arrival, index, spec)
        active = {}                     # job id -> (arrival, index, spec, bytes)
        lat, done_samples, checks, check_futs = [], 0, [], []
        n_done = 0
        t0 = time.perf_counter()
        steps = 0
        step_ms = 0.0
        max_active = 0
        t_report = 0.0
        t_last = requests[-1][0] if requests else 0.0
        unfinished = 0
        while reqs or pending or active or self.building:
            now = time.perf_counter() - t0
            if drain_limit is not None and not reqs and now > t_last + drain_limit:
                unfinished = len(pending) + len(active)
                break
            if log and now - t_report >= 10.0:
                t_report = now
                log(f"serving t={now:.0f}s: {len(lat)} done, {len(active)} active, "
                    f"{len(pending)} waiting, {len(self.scenes)} scenes resident")
            while reqs and reqs[0][0] <= now:
                pending.append(reqs.popleft())
            # finished scene builds
            for spec, fut in list(self.building.items()):
                if fut.done():
                    sd, sc = fut.result()
                    del self.building[spec]
                    self.descs[spec] = sd
                    self.scenes[spec] = sc
                    fp = sc.footprint()
                    if not self.planner.resize_scene(spec, fp):
                        # measured above the estimate: evict idle scenes
                        ev = self.planner.make_room(fp - self.planner.scenes[spec][0],
                                                    protect={spec} | set(self.building))
                        if ev is None:
                            raise RuntimeError("scene does not fit the HBM budget")
                        for k in ev:
                            self._evict(k)
                        assert self.planner.resize_scene(spec, fp)
            # admission, in arrival order; a job waiting for its scene does
            # not block later jobs whose scenes are resident
            keep = collections.deque()
            blocked_mem = False
            starving = set()    # scenes waiting jobs need that find no room
            while pending:
                a, idx, job = pending.popleft()
                spec = job.scene
                nbytes = native.Engine.footprint(job.params())
                if spec in self.scenes and spec not in self.draining and not blocked_mem:
                    if self.planner.can_admit(spec, nbytes):
                        self.planner.admit(spec, nbytes)
                        jid = self.engine.submit(self.scenes[spec], job.camera(), job.params())
                        active[jid] = (a, idx, job, nbytes)
                        continue
                    blocked_mem = True
                elif spec not in self.building and spec not in self.scenes:
                    if spec.estimated_bytes() + nbytes > self.planner.budget - self.planner.reserved:
                        raise RuntimeError(f"job {idx}: scene and accumulators exceed the budget")
                    # scenes being built or wanted by waiting jobs stay
                    wanted = set(self.building) | {j.scene for _, _, j in pending}
                    wanted |= {j.scene for _, _, j in keep}
                    wanted -= self.draining
                    ev = self.planner.plan_scene(spec, spec.estimated_bytes(), protect=wanted)
                    if ev is not None:
                        for k in ev:
                            self._evict(k)
                        self.planner.add_scene(spec, spec.estimated_bytes())
                        self.building[spec] = self.builder.submit(self._build, spec)
                    else:
                        starving.add(spec)
                keep.append((a, idx, job))
            pending = keep
            # Without this, scenes whose jobs keep arriving never go idle
            # and a job on a scene that is not resident waits for ever.
            # Once fewer scenes starve, the most wanted drained scenes take
            # jobs again.
            need = self.planner.drain_target(len(starving)) - len(self.draining)
            if need:
                demand = collections.Counter(j.scene for _, _, j in pending)
                if need > 0:
                    self.draining.update(self.planner.drain_victims(
                        need, demand, self.draining, protect=set(self.building)))
                else:
                    back = sorted(self.draining, key=lambda k: -demand.get(k, 0))[:-need]
                    self.draining.difference_update(back)
            max_active = max(max_active, len(active))
            if active:
                si = self.engine.step()
                steps += 1
                step_ms += si["ms_device"]
                if si["jobs_finished"]:
                    tnow = time.perf_counter() - t0
                    for jid in list(active):
                        st = self.engine.poll(jid)
                        if st["state"] != abi.SGR_JOB_DONE:
                            continue
                        a, idx, job, nbytes = active.pop(jid)
                        lat.append(tnow - a)
                        done_samples += job.samples
                        n_done += 1
                        acc = None
                        if keep_results:
                            acc = results[idx] = self.engine.result(jid, job.width, job.height)
                        if self.check_every and n_done % self.check_every == 0:
                            if acc is None:
                                acc = self.engine.result(jid, job.width, job.height)
                            check_futs.append((idx, self.checker.submit(
                                self.check_fn, job, self.descs[job.scene], acc)))
                        self.engine.release(jid)
                        self.planner.release(job.scene, nbytes)
            else:
                wait = 0.001
                if reqs and not pending and not self.building:
                    wait = max(0.0, min(0.05, reqs[0][0] - (time.perf_counter() - t0)))
                time.sleep(wait)
            if not drain and not reqs:
                break
        wall = time.perf_counter() - t0
        for idx, f in check_futs:
            checks.append(dict(f.result(), job=idx))
        lat_ms = np.array(lat) * 1e3
        return {
            "jobs": len(lat), "wall_s": wall,
            "samples": int(done_samples),
            "samples_per_s": done_samples / wall if wall > 0 else 0.0,
            "p50_ms": float(np.percentile(lat_ms, 50)) if len(lat) else None,
            "p99_ms": float(np.percentile(lat_ms, 99)) if len(lat) else None,
            "mean_ms": float(lat_ms.mean()) if len(lat) else None,
            "steps": steps, "device_ms": step_ms,
            "max_active_jobs": max_active,
            "peak_budget_bytes": self.planner.peak, "budget_bytes": self.planner.budget,
            "resident_scenes_max": self.planner.max_scenes,
            "checks": checks,
            "unfinished": unfinished,
            **({"results": results} if keep_results else {}),
        }

    def _evict(self, spec):
        self.planner.evict(spec)
        self.draining.discard(spec)
        sc = self.scenes.pop(spec, None)
        if sc is not None:
            sc.close()
        self.descs.pop(spec, None)

    def close(self):
        self.builder.shutdown(wait=True)
        self.checker.shutdown(wait=True)
        self.engine.release_all()
        for spec in list(self.scenes):
            self.scenes.pop(spec).close()


def serve_poisson(ctx, rate, duration, n_scene_pool=48, seed=1, budget=None, log=None,
                  prewarm=True, drain_limit=None, **kw):
    """Poisson arrivals at RATE jobs/s for DURATION s, jobs drawn from the
    batched-job distribution over a pool of scenes.  With PREWARM the
    first max_scenes scenes of the pool are made resident before the
    first arrival (the steady state of a long-running server)."""
    from . import workloads
    pool = workloads.scene_pool(n_scene_pool, seed)
    times = workloads.poisson_arrivals(rate, duration, seed + 1)
    jobs = workloads.draw_jobs(len(times), seed + 2, scene_pool=pool)
    if budget is None:
        budget = 64 << 30
    srv = Server(ctx, budget, **kw)
    try:
        t_warm = srv.prewarm(pool) if prewarm else 0.0
        rep = srv.run(list(zip(times, jobs)), log=log, drain_limit=drain_limit)
    finally:
        srv.close()
    rep.update({"rate_jobs_per_s": rate, "duration_s": duration, "scene_pool": n_scene_pool,
                "prewarm_s": t_warm})
    return rep
