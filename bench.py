"""Headline benchmark: plain GPU path tracing of the seeded 1.31M-triangle
icosphere soup at 1920x1080, 64 spp, max path length 10, on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]

One step is one full render (every pixel, every sample, filter margin
included) into device-resident accumulators; the scene is built and
uploaded before the timed region.  Prints one JSON line.
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "samples/s at 1080p; equal-time relMSE vs oracle; p50/p99 job latency"
W, H, SPP, MAXP, MINP = 1920, 1080, 64, 10, 3
HBM_PEAK_GBS = 8000.0
# Algorithmic bytes of one closest-hit / any-hit query, from the counted
# work: a 16-byte octree node per node visit, a 36-byte triangle (corner
# and two edges) plus its 4-byte surface-list entry per primitive test,
# and the ray record (32 B in; 16 B hit out / 1 B flag out + 4 B queue id).
NODE_B, TEST_B, LIST_B = 16, 36, 4
RAY_IN_B, HIT_OUT_B, OCC_OUT_B, QUEUE_B = 32, 16, 1, 4


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def algorithmic_bytes(st):
    ext = (st["node_visits_closest"] * NODE_B
           + st["surf_tests_closest"] * (TEST_B + LIST_B)
           + st["scene_intersect_calls"] * (RAY_IN_B + HIT_OUT_B + QUEUE_B))
    occ = (st["node_visits_shadow"] * NODE_B
           + st["surf_tests_shadow"] * (TEST_B + LIST_B)
           + st["shadow_tests"] * (RAY_IN_B + OCC_OUT_B + QUEUE_B))
    return ext, occ


def read_traffic(kernel):
    """HBM bytes per launch measured by rocprofv3 --pmc (FETCH_SIZE +
    WRITE_SIZE, scaled by the copy-kernel calibration), if a committed
    summary records it."""
    path = os.path.join(ROOT, "profiles", "pmc_traffic_icosphere_soup_1080p.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith(f"traffic_bytes_per_launch {kernel} "):
            return float(line.split()[2])
    return None


def cpu_baseline(sd, cam, seed):
    """The CPU oracle (the reference's algorithm restated in C++) on a
    bounded crop of the same workload: the centred 480x270 window of the
    1920x1080 image at 16 spp, all allowed host cores."""
    import oracle
    try:
        cores = min(len(os.sched_getaffinity(0)), 16)
    except AttributeError:
        cores = min(os.cpu_count() or 1, 16)
    from snogray_amd import abi
    osc = oracle.OracleScene(sd)
    ww, hh, spp = 480, 270, 16
    p = abi.make_params(W, H, spp, max_path_len=MAXP, min_path_len=MINP, seed=seed,
                        window=((W - ww) // 2, (H - hh) // 2, ww, hh))
    t = time.perf_counter()
    _, _, st = osc.render(cam, p, nthreads=cores)
    dt = time.perf_counter() - t
    return {"value": ww * hh * spp / dt, "unit": "samples/s", "cores": cores,
            "kind": "port",
            "sample": f"centred {ww}x{hh} window of the 1920x1080 image, {spp} spp, "
                      f"max_path_len {MAXP}, same scene and camera ({dt:.1f} s)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--calibrate", action="store_true",
                    help="also run one plain 256 MiB device copy (reference kernel "
                         "for scaling rocprofv3 FETCH_SIZE / WRITE_SIZE)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); "
              "the multi-GPU sharding is a design in DESIGN.md, not a measured path.",
              file=sys.stderr)
        sys.exit(2)

    import torch
    from snogray_amd import abi, native, scenes

    log("building scene")
    t0 = time.perf_counter()
    sd, cam = scenes.icosphere_soup()
    t_gen = time.perf_counter() - t0
    ctx = native.Context(0)
    t0 = time.perf_counter()
    sc = native.Scene(ctx, sd)
    t_build = time.perf_counter() - t0
    acc = torch.zeros((H, W, 5), dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    p = abi.make_params(W, H, SPP, max_path_len=MAXP, min_path_len=MINP, seed=args.seed,
                        flags=abi.SGR_RENDER_KERNEL_TIMING)
    for i in range(args.warmup):
        log(f"warmup {i}")
        sc.render_device(cam, p, acc.data_ptr())
    torch.cuda.synchronize()
    tot = None
    t0 = time.perf_counter()
    for i in range(args.steps):
        st = sc.render_device(cam, p, acc.data_ptr())
        log(f"step {i}: {st['ms_render']:.1f} ms")
        if tot is None:
            tot = dict(st)
        else:
            for k, v in st.items():
                tot[k] += v
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_step = 1e3 * wall / args.steps
    if args.calibrate:
        src = torch.ones(64 << 20, dtype=torch.float32, device="cuda")
        dst = src.clone()
        torch.cuda.synchronize()
        log(f"calibration copy: {src.numel() * 4} bytes read, {dst.numel() * 4} written")
    samples = W * H * SPP
    value = samples * args.steps / wall
    assert torch.isfinite(acc).all().item()

    ext_b, occ_b = algorithmic_bytes(tot)
    kernels = {"k_extend": (tot["ms_extend"], ext_b, tot["extend_launches"]),
               "k_occlude": (tot["ms_shadow"], occ_b, tot["shadow_launches"])}
    dom = max(kernels, key=lambda k: kernels[k][0])
    ms, bytes_, launches = kernels[dom]
    achieved = bytes_ / (ms / 1e3) / 1e9
    traffic = read_traffic(dom)
    roofline = {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 2),
                "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": traffic,
                "algorithmic_bytes_per_launch": bytes_ / launches,
                "avg_launch_ms": ms / launches}
    cpu = None
    if not args.no_cpu:
        log("cpu baseline")
        cpu = cpu_baseline(sd, cam, args.seed)
    per_step = {k: tot[k] / args.steps for k in tot}
    line = {
        "metric": METRIC, "value": round(value, 1), "unit": "samples/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_step, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: seeded icosphere soup and Gaussian-lobe envmap built in code",
        "config": {"workload": "icosphere soup, 1920x1080, 64 spp, max path length 10",
                   "triangles": int(sd.num_triangles), "width": W, "height": H,
                   "spp": SPP, "max_path_len": MAXP, "min_path_len": MINP,
                   "envmap": "1024x512 fp32", "parallelism": "single GPU"},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "gpu_vs_cpu": round(value / cpu["value"], 1) if cpu else None,
        "end_to_end_samples_per_s": round(samples / (ms_step / 1e3 + t_build), 1),
        "scene_build_s": round(t_build, 3), "scene_gen_s": round(t_gen, 3),
        "kernel_ms_per_step": {k: round(per_step[k], 3) for k in
                               ("ms_extend", "ms_shade", "ms_shadow", "ms_splat", "ms_other")},
        "counters_per_step": {k: int(per_step[k]) for k in
                              ("scene_intersect_calls", "shadow_tests", "node_visits_closest",
                               "surf_tests_closest", "node_visits_shadow",
                               "surf_tests_shadow")},
    }
    print(json.dumps(line), flush=True)
    sc.close()
    ctx.close()


if __name__ == "__main__":
    main()
