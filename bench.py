#!/usr/bin/env python3
"""Benchmark: GPIS path tracing throughput on one MI355X (Mpaths/s).

Workload (one step): the sphere-mean GPIS scene in renewal_plus mode at
256x256, 64 spp, max_bounce 10, kernel sigma 1.5 / length_scale 2 / aniso
[1,1,4] / 3 points per cell, seeded 64x32 environment map -- every pass of
every pixel, i.e. 4,194,304 paths, rendered into the FP64 accumulator.

Reported next to it:
  roofline      the GP-medium march kernel (k_march_cached<ppc>): algorithmic FP32
                flops per launch (field evaluations x (27*ppc*16 + 9)) over its
                HIP-event time, against the MI355X FP32 vector peak; its HBM traffic
                per launch is the rocprofv3 FETCH_SIZE + WRITE_SIZE figure measured for
                this same workload, scaled by a device copy of known size;
  cpu_baseline  the CPU restatement of the reference trace loop (OpenMP over
                pixels, static schedule) on a bounded sample of the same
                workload, timed on this host's cores.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

FP32_PEAK_TFLOPS = 157.3
# HBM bytes per k_march_cached<3> launch on this workload: rocprofv3 --pmc FETCH_SIZE and
# WRITE_SIZE over `bench.py --steps 1 --warmup 0 --copy-probe-mb 256`, FETCH_SIZE scaled by
# 2.0 and WRITE_SIZE by 1.0 (what the counters read for the 256 MiB copy);
# profiles/k_march_cached_sphere_renewal_256x256x64.md holds the table.
MARCH_TRAFFIC_BYTES = 101_084_731
CFG = 2
WORKLOAD = ("sphere-mean GPIS, renewal_plus, 256x256, 64 spp, max_bounce 10, "
            "kernel sigma 1.5 / length_scale 2 / aniso [1,1,4] / ppc 3, "
            "mean sphere r70 @ (0,0,-500) in a null r130 sphere, seeded 64x32 env map")


def cpu_baseline(scene: dict, target_s: float = 15.0) -> dict:
    """Time the oracle's faithful restatement on a bounded, deterministic pixel sample."""
    from oracle import oracle as O
    osc = O.OracleScene(scene)
    threads = O.lib().orc_max_threads()
    W, H = scene["width"], scene["height"]
    stride = 64
    _, paths, secs = osc.render(0, 1, pixel_stride=stride, pixel_offset=7, nthreads=threads)
    # grow the sample toward target_s of CPU time (every sample is a fixed pixel subset, pass 0)
    if secs < target_s / 4:
        stride = max(1, int(stride * secs / (target_s / 2))) if secs > 0 else 8
        _, paths, secs = osc.render(0, 1, pixel_stride=stride, pixel_offset=7 % stride,
                                    nthreads=threads)
    return {
        "value": paths / secs / 1e6, "unit": "Mpaths/s", "cores": threads, "kind": "port",
        "sample": f"pass k=0 of every {stride}th pixel ({paths} paths of the {W}x{H} image), "
                  f"{secs:.1f} s wall on {threads} OpenMP threads",
    }


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--copy-probe-mb", type=int, default=0,
                    help="copy a tensor of this many MiB before the timed region, so that a "
                         "profiler's FETCH_SIZE / WRITE_SIZE can be scaled by a known traffic")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the GPIS path tracer shards by pixel tiles or passes "
              "with a single FP64 image reduce, which is not part of this build; run --gpus 1.",
              file=sys.stderr)
        return 2

    import torch
    from mupsi_amd import Context, Scene, scenes

    scene = scenes.config_scene(CFG)
    W, H, spp = scene["width"], scene["height"], scene["spp"]
    ctx = Context(0)
    s = Scene(ctx, scene)
    print(f"bench: {ctx.info()} scene {W}x{H} spp {spp}", file=sys.stderr, flush=True)

    if args.copy_probe_mb:
        src = torch.empty(args.copy_probe_mb << 18, dtype=torch.float32, device="cuda").fill_(1.0)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        print(f"bench: copy probe of {args.copy_probe_mb} MiB done", file=sys.stderr, flush=True)
        del src, dst

    for i in range(args.warmup):
        s.reset()
        s.render_passes(0, spp)
        print(f"bench: warmup {i} done", file=sys.stderr, flush=True)

    march_ms = 0.0
    launches = 0
    evals = 0
    steps_ref = 0
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        s.reset()
        s.render_passes(0, spp)
        st = s.stats()
        march_ms += st["march_kernel_ms"]
        launches += st["march_kernel_launches"]
        evals += st["field_evals"]
        steps_ref += st["march_steps"]
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    secs = t1 - t0
    paths = W * H * spp * args.steps
    ms_per_step = 1e3 * secs / args.steps
    ppc = scene["kernel"]["points_per_cell"]
    flops_per_eval = 27 * ppc * 16 + 9
    achieved = evals * flops_per_eval / (march_ms * 1e-3) / 1e12 if march_ms > 0 else 0.0
    img = s.image()
    print(f"bench: {args.steps} steps in {secs:.3f} s; march kernel {march_ms:.1f} ms over "
          f"{launches} launches; {evals} field evals (reference grid points {steps_ref}); "
          f"image mean {img.mean():.5f}", file=sys.stderr, flush=True)

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(scene)
        print(f"bench: cpu baseline {cpu}", file=sys.stderr, flush=True)

    line = {
        "metric": "Mpaths/s + time-to-1e-3 relMSE: guided vs plain GPU tracer, 1 MI355X",
        "value": paths / secs / 1e6,
        "unit": "Mpaths/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": WORKLOAD, "width": W, "height": H, "spp": spp,
                   "paths_per_step": W * H * spp, "tracer": "plain wavefront"},
        "roofline": {
            "bound": "valu",
            "kernel": f"k_march_cached<{ppc}>",
            "achieved": achieved,
            "peak": FP32_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / FP32_PEAK_TFLOPS,
            "traffic": MARCH_TRAFFIC_BYTES,
            "flops_per_launch": evals * flops_per_eval / max(1, launches),
            "avg_launch_ms": march_ms / max(1, launches),
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    s.close()
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
