#!/usr/bin/env python3
"""Benchmark: GPIS path tracing throughput on one MI355X (Mpaths/s).

Workload (one step): the sphere-mean GPIS scene in renewal_plus mode at
256x256, 64 spp, max_bounce 10, kernel sigma 1.5 / length_scale 2 / aniso
[1,1,4] / 3 points per cell, seeded 64x32 environment map -- every pass of
every pixel, i.e. 4,194,304 paths, rendered into the FP64 accumulator.

Reported next to it:
  roofline      the GP-medium march kernel (k_march_sphere<ppc, 16>): algorithmic FP32
                flops per launch (field evaluations x (27*ppc*16 + 9)) over its
                HIP-event time, against the MI355X FP32 vector peak; its HBM traffic
                per launch is the rocprofv3 FETCH_SIZE + WRITE_SIZE figure measured for
                this same workload, scaled by a device copy of known size;
  cpu_baseline  the CPU restatement of the reference trace loop (OpenMP over
                pixels, static schedule) on a bounded sample of the same
                workload, timed on this host's cores.
  guiding       the learned-guiding workload (the same sphere scene with a
                mirror sphere and an emissive quad, 1024x1024): wall time for
                the plain and the guided tracer to reach 1e-3 relative MSE
                (mean over pixels and channels of (I - R)^2 / (R^2 + 0.01))
                against a 6144-pass plain R, measured on progressive renders
                (the guided time includes its 32 online-training passes), with
                the figure extrapolated from per-pass variances beside it.
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
# HBM bytes per k_march_sphere<3, 16> launch on this workload (the step's 9 launches averaged):
# rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE over `bench.py --steps 1 --warmup 0 --copy-probe-mb 256`,
# FETCH_SIZE scaled by 2.0 and WRITE_SIZE by 1.0 (what the counters read for the 256 MiB copy);
# profiles/k_march_sphere_sphere_renewal_256x256x64.md holds the table.
MARCH_TRAFFIC_BYTES = 112_700_000
PRESET = "sphere_renewal_256"
WORKLOAD = ("sphere-mean GPIS, renewal_plus, 256x256, 64 spp, max_bounce 10, "
            "kernel sigma 1.5 / length_scale 2 / aniso [1,1,4] / ppc 3, "
            "mean sphere r70 @ (0,0,-500) in a null r130 sphere, seeded 64x32 env map")


ORACLE_FLAGS = "g++ -O3 -std=c++20 -fopenmp -ffp-contract=off -mavx2 -mfma"


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(scene: dict, target_s: float = 15.0) -> dict:
    """Time the oracle's faithful restatement on a bounded, deterministic pixel sample.

    The oracle is built with ORACLE_FLAGS (oracle/Makefile): x86-64-v3 instead of
    the reference's -march=native, because the library is built in one place and
    timed on another host, and -ffp-contract=off so that its float operation
    order is the documented one."""
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
        "cpu_model": cpu_model(), "build": ORACLE_FLAGS,
    }


def _chunk_stats(s, k0: int, chunks: int, spp: int, R):
    """Per-pass relative variance v (mean over pixels and channels of the
    per-pixel variance of one pass over R^2 + 0.01, R an independent reference
    image) and seconds per pass, from `chunks` independent renders of `spp`
    passes.  R must not come from the same samples: a pixel whose samples hold a
    rare bright path would otherwise carry it in its own denominator."""
    import numpy as np
    import torch
    ims = []
    secs = 0.0
    for c in range(chunks):
        s.reset()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        s.render_passes(k0 + c * spp, spp)
        torch.cuda.synchronize()
        secs += time.perf_counter() - t0
        ims.append(s.image())
    ims = np.array(ims)
    v = float(np.mean(ims.var(0, ddof=1) * spp / (R ** 2 + 0.01)))
    return v, secs / (chunks * spp)


def _relmse(img, R) -> float:
    import numpy as np
    return float(np.mean((img - R) ** 2 / (R ** 2 + 0.01)))


def _crossing(s, R, k0: int, target: float, first: int, step: int, limit: int,
              r_err: float) -> dict:
    """Progressive render from an empty image until its relMSE against R, less
    R's own estimated relMSE r_err (the two are independent), falls to target:
    passes are added `first` then `step` at a time and the image is read back
    between calls (outside the timed render calls).  Returns the pass count
    and the summed wall time of the render calls at the first check at or
    below target, the raw relMSE there, and the history of checks."""
    import torch
    s.reset()
    n, secs, hist = 0, 0.0, []
    while n < limit:
        k = first if n == 0 else step
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        s.render_passes(k0 + n, k)
        torch.cuda.synchronize()
        secs += time.perf_counter() - t0
        n += k
        e = _relmse(s.image(), R)
        hist.append((n, round(secs, 4), e))
        if e - r_err <= target:
            return {"passes": n, "time_s": secs, "relmse_raw": e, "relmse_less_reference": e - r_err,
                    "crossed": True, "checks": len(hist)}
    return {"passes": n, "time_s": secs, "relmse_raw": hist[-1][2], "crossed": False,
            "checks": len(hist)}


def guiding_bench(ctx, target: float = 1e-3, chunks: int = 8, spp: int = 8,
                  ref_passes: int = 6144) -> dict:
    """Time to `target` relMSE, plain vs guided, on the learned-guiding workload.

    relMSE(I) = mean over pixels and channels of (I - R)^2 / (R^2 + 0.01).  R is
    a ref_passes-pass plain render on its own pass range (its relMSE, v_p /
    ref_passes, is about 1e-4 and is subtracted from the measured figures).
    Measured: the plain tracer and the guided tracers (the network trained
    online on its first 32 passes, and on its first 16; and the network with
    the reference's Russian roulette, rr_boost 1) render progressively until
    their image crosses target; the wall time of the render calls up to that
    pass is the time to target (training passes and Adam steps included).
    Extrapolated beside it: both tracers are unbiased, so an N-pass image has
    relMSE v / N with v the per-pass variance over R^2 + 0.01, estimated from
    `chunks` independent renders of `spp` passes; time = seconds per pass x
    v / target."""
    import numpy as np
    import torch
    from mupsi_amd import Guide, Scene, scenes
    sc = scenes.preset_scene("guided_sphere_quad_1024")
    W, H = sc["width"], sc["height"]
    plain = Scene(ctx, sc)
    plain.render_passes(0, 1)                         # allocations outside the timing
    plain.reset()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    plain.render_passes(100000, ref_passes)
    torch.cuda.synchronize()
    ref_s = time.perf_counter() - t0
    R = plain.image()
    print(f"bench: guiding reference {ref_passes} passes in {ref_s:.1f} s", file=sys.stderr, flush=True)
    v_p, t_p = _chunk_stats(plain, 1000, chunks, spp, R)
    r_err = v_p / ref_passes
    plain_x = _crossing(plain, R, 0, target, 64, 16, 4096, r_err)
    print(f"bench: guiding plain crossing {plain_x}", file=sys.stderr, flush=True)
    plain.close()
    guided = Scene(ctx, sc)
    out_runs = {}
    for name, cfg in (("guided_train32", dict(train_passes=32)),
                      ("guided_train16", dict(train_passes=16, passes_per_update=2)),
                      ("network_only_reference_roulette", dict(train_passes=32, rr_boost=1.0))):
        g = Guide.for_scene(ctx, sc, seed=7, **cfg)
        n_t = g.config["train_passes"]
        if not out_runs:   # allocate the path, guide and ring buffers outside the timing
            warm = Guide.for_scene(ctx, sc, seed=99, **cfg)
            guided.set_guide(warm)
            guided.render_passes(50000, n_t + 4)
            warm.close()
        guided.set_guide(g)                            # fresh network and training history
        x = _crossing(guided, R, 2000 + 1000 * len(out_runs), target, n_t + 8, 4, 4096, r_err)
        st = g.stats()
        x.update({"train_passes": n_t, "rr_boost": g.config["rr_boost"],
                  "train_steps": st["train_steps"], "train_ms": st["train_ms"],
                  "fraction_of_plain_time": x["time_s"] / plain_x["time_s"]})
        if name == "guided_train32":   # the trained network's per-pass variance and speed
            v_g, t_g = _chunk_stats(guided, 3000, chunks, spp, R)
            x["v_per_pass_trained"] = v_g
            x["s_per_pass_trained"] = t_g
            mlp = {"steps": st["train_steps"], "ms": st["train_ms"]}
        out_runs[name] = x
        print(f"bench: guiding {name} {x}", file=sys.stderr, flush=True)
        g.close()
    # extrapolated (per-pass variance x seconds per pass), as a cross-check
    gx = out_runs["guided_train32"]
    time_plain_x = t_p * v_p / target
    n_t = gx["train_passes"]
    train_s = gx["time_s"] - (gx["passes"] - n_t) * gx["s_per_pass_trained"]
    time_guided_x = train_s + gx["s_per_pass_trained"] * gx["v_per_pass_trained"] / target
    # the guided tracer's roulette without the network (alpha = 0)
    ctl = Guide.for_scene(ctx, sc, alpha=0.0, train_passes=0)
    guided.set_guide(ctl)
    v_c, _ = _chunk_stats(guided, 4000, chunks // 2, spp, R)
    ctl.close()
    guided.close()
    paths = W * H
    # algorithmic GEMM flops per training sample: forward 2(64*32 + 2*64*64 + 41*64),
    # input gradients of layers 4..2 2(41*64 + 2*64*64), weight gradients as forward
    mflop_step = 73088 * Guide.DEFAULTS["batch_size"] / 1e6
    return {
        "workload": f"sphere-mean GPIS renewal_plus + mirror sphere r0.5 + emissive quad "
                    f"(area 0.01, radiance 200), {W}x{H}",
        "target_relmse": target,
        "reference": {"passes": ref_passes, "s": ref_s, "relmse_of_reference": r_err},
        "measured": {"plain": plain_x, **out_runs},
        "guided_time_fraction": gx["fraction_of_plain_time"],
        "network_only_time_fraction": out_runs["network_only_reference_roulette"]["fraction_of_plain_time"],
        "guided_train16_time_fraction": out_runs["guided_train16"]["fraction_of_plain_time"],
        "extrapolated": {
            "plain": {"v_per_pass": v_p, "s_per_pass": t_p, "Mpaths_s": paths / t_p / 1e6,
                      "time_to_target_s": time_plain_x},
            "guided_train32": {"v_per_pass": gx["v_per_pass_trained"],
                               "s_per_pass": gx["s_per_pass_trained"],
                               "Mpaths_s": paths / gx["s_per_pass_trained"] / 1e6,
                               "time_to_target_s": time_guided_x},
            "guided_time_fraction": time_guided_x / time_plain_x,
            "estimates": f"{chunks} x {spp} passes per chunk"},
        "decomposition_v_per_pass": {"plain": v_p, "roulette_only_alpha0": v_c,
                                     "guided": gx["v_per_pass_trained"]},
        "mlp_train": {"steps": mlp["steps"], "ms": mlp["ms"],
                      "TFLOP_s": (mlp["steps"] * mflop_step / 1e6) / (mlp["ms"] * 1e-3)
                      if mlp["ms"] > 0 else 0.0,
                      "peak_bf16_TFLOP_s": 2500.0},
    }


HBM_PEAK_GBS = 8000.0


def traversal_bench(ctx, sc: dict, n: int = 1 << 22) -> dict:
    """BVH traversal memory roofline on the mean mesh of `sc`.

    n query points in a shell around the surface (a random vertex scaled by
    U[0.75, 1.25], then the mean's placement): the region where a mesh-mean
    march needs exact distance and inside queries.  Algorithmic bytes per query
    are 32 per node whose bounds are tested plus 52 per triangle tested (index,
    face, three vertices), counted by the same traversal
    (mupsi_mesh_query_cost); the kernel time is the HIP-event time of the query
    kernel on the library's stream (median of 3).  The closest-point search
    starts from the distance grid's bound (as the renderer runs it); the
    unbounded search's counts are reported beside it."""
    import numpy as np
    import torch
    from mupsi_amd import Scene, batch
    s = Scene(ctx, sc)
    m = sc["gp_medium"]["mesh"]
    xf = sc["gp_medium"]["transform"]
    rng = np.random.default_rng(3)
    v = m["vertices"][rng.integers(0, len(m["vertices"]), n)].astype(np.float64)
    v *= rng.uniform(0.75, 1.25, (n, 1))
    p = (np.asarray(xf["position"]) + xf["scale"] * v).astype(np.float32)
    pt = torch.from_numpy(p).cuda()
    out = {"queries": n, "triangles": int(len(m["faces"])),
           "points": "random vertex x U[0.75, 1.25], world space"}
    n0, t0 = batch.mesh_query_cost(s, -1, 0, pt)
    out["closest_unbounded_nodes_per_query"] = float(n0.sum().item()) / n
    out["closest_unbounded_tris_per_query"] = float(t0.sum().item()) / n
    # inside: the reference's loop of casts (kind 1) and the stackless walk over the
    # tree (kind 4) beside the pass over the point's leaf bin the kernel runs
    # (kind 3: a "node" is a bin entry, 32 bytes like a node), whose counts give
    # the algorithmic bytes
    n1, t1 = batch.mesh_query_cost(s, -1, 1, pt)
    out["inside_reference_casts_nodes_per_query"] = float(n1.sum().item()) / n
    out["inside_reference_casts_tris_per_query"] = float(t1.sum().item()) / n
    n4, t4 = batch.mesh_query_cost(s, -1, 4, pt)
    out["inside_walk_nodes_per_query"] = float(n4.sum().item()) / n
    out["inside_walk_tris_per_query"] = float(t4.sum().item()) / n
    for kind, name, fn in ((2, "closest", batch.mesh_closest), (3, "inside", batch.mesh_inside)):
        nodes, tris = batch.mesh_query_cost(s, -1, kind, pt)
        nb = int(nodes.sum().item())
        tb = int(tris.sum().item())
        algo = 32 * nb + 52 * tb
        fn(s, -1, pt)
        ms = []
        for _ in range(3):
            fn(s, -1, pt)
            ms.append(batch.last_kernel_ms(ctx))
        t = float(np.median(ms))
        gbs = algo / (t * 1e-3) / 1e9
        out[name] = {"kernel": "k_mesh_closest" if kind == 2 else "k_mesh_inside",
                     "kernel_ms": t, "Mqueries_s": n / t / 1e3,
                     "nodes_per_query": nb / n, "tris_per_query": tb / n,
                     "algorithmic_bytes": algo, "GB_s": gbs,
                     "frac_of_hbm_peak": gbs / HBM_PEAK_GBS}
    s.close()
    return out


def configs_bench(ctx, mesh_passes: int = 64, batch_div: int = 16) -> dict:
    """Throughput on the other workloads: sphere_single_128 (single_realization,
    128x128x16, every pass), icosphere_renewal_512 (mesh mean, 512x512, all 64
    passes by default) and
    generated_batch (64 generated scenes, the first spp/batch_div passes of each,
    rendered scene by scene and as one autotuned batch; the two must give
    identical images)."""
    import numpy as np
    import torch
    from mupsi_amd import Batch, Scene, scenes
    out = {}

    def timed(fn):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        fn()
        torch.cuda.synchronize()
        return time.perf_counter() - t0

    sc1 = scenes.preset_scene("sphere_single_128")
    s1 = Scene(ctx, sc1)
    # the whole 262,144-path render is a few milliseconds: its bounce loop replays a
    # captured graph (as the headline render does), captured by a warm-up render of
    # the same size; the median of 5 timed renders
    s1.set_graph(True)
    s1.render_passes(0, sc1["spp"])
    n1 = sc1["width"] * sc1["height"] * sc1["spp"]
    times = []
    for _ in range(5):
        s1.reset()
        times.append(timed(lambda: s1.render_passes(0, sc1["spp"])))
    secs = float(np.median(times))
    out["sphere_single_128"] = {"workload": "sphere mean, single_realization, 128x128, 16 spp",
                   "paths": n1, "s": secs, "Mpaths_s": n1 / secs / 1e6,
                   "bounce_loop": "hipGraph replay", "timed_renders": 5}
    s1.close()
    print(f"bench: sphere_single_128 {out['sphere_single_128']}", file=sys.stderr, flush=True)

    sc3 = scenes.preset_scene("icosphere_renewal_512")
    t0 = time.perf_counter()
    s3 = Scene(ctx, sc3)
    upload = time.perf_counter() - t0
    s3.render_passes(1000, 1)
    s3.reset()
    n3 = sc3["width"] * sc3["height"] * mesh_passes
    secs = timed(lambda: s3.render_passes(0, mesh_passes))
    st = s3.stats()
    out["icosphere_renewal_512"] = {"workload": "displaced icosphere mean (327,680 triangles), renewal_plus, "
                               f"512x512, passes 0..{mesh_passes - 1} of 64",
                   "paths": n3, "s": secs, "Mpaths_s": n3 / secs / 1e6,
                   "upload_s": upload, "march_kernel_ms": st["march_kernel_ms"],
                   "field_evals": st["field_evals"]}
    s3.close()
    print(f"bench: icosphere_renewal_512 {out['icosphere_renewal_512']}", file=sys.stderr, flush=True)
    out["bvh_traversal"] = traversal_bench(ctx, sc3)
    print(f"bench: bvh {out['bvh_traversal']}", file=sys.stderr, flush=True)

    scs = scenes.generated_batch()
    kc = [max(1, sc["spp"] // batch_div) for sc in scs]
    n4 = sum(sc["width"] * sc["height"] * k for sc, k in zip(scs, kc))
    t0 = time.perf_counter()
    seq = [Scene(ctx, sc) for sc in scs]
    upload = time.perf_counter() - t0
    for x in seq:
        x.render_passes(100, 1)
        x.reset()
    secs_seq = timed(lambda: [x.render_passes(0, k) for x, k in zip(seq, kc)])
    ref = [x.image() for x in seq]
    for x in seq:
        x.reset()
    b = Batch(ctx, seq)
    secs_b = timed(lambda: b.render([0] * len(scs), kc))
    same = all(np.array_equal(r, x.image()) for r, x in zip(ref, seq))
    bst = b.stats()
    tun = b.tuning()
    b.close()
    for x in seq:
        x.close()
    out["generated_batch"] = {"workload": f"64 generated scenes (sphere or icosphere means, kr logU[0.05,5], "
                               f"ppc 1..8, 64..512 px), the first spp/{batch_div} passes of each",
                   "paths": n4, "upload_s": upload,
                   "sequential": {"s": secs_seq, "Mpaths_s": n4 / secs_seq / 1e6},
                   "batched": {"s": secs_b, "Mpaths_s": n4 / secs_b / 1e6,
                               "launches": bst["launches"], "tune_ms": bst["tune_ms"],
                               "paths_per_launch": tun["paths_per_launch"],
                               "march_blocks_per_cu": tun["march_blocks_per_cu"],
                               "use_graph": tun["use_graph"]},
                   "batched_equals_sequential": bool(same)}
    print(f"bench: generated_batch {out['generated_batch']}", file=sys.stderr, flush=True)
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-graph", action="store_true",
                    help="issue the bounce loop's launches one by one instead of replaying a graph")
    ap.add_argument("--no-guiding", action="store_true", help="skip the guided-vs-plain measurement")
    ap.add_argument("--no-configs", action="store_true",
                    help="skip the sphere_single_128, icosphere_renewal_512, generated_batch and BVH traversal measurements")
    ap.add_argument("--only", choices=["guiding", "configs", "bvh"], default=None,
                    help="run one section alone and print its JSON (for profiling runs)")
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

    def copy_probe(mb: int) -> None:
        src = torch.empty(mb << 18, dtype=torch.float32, device="cuda").fill_(1.0)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        print(f"bench: copy probe of {mb} MiB done", file=sys.stderr, flush=True)
        del src, dst

    if args.only:
        ctx = Context(0)
        print(f"bench: {ctx.info()} section {args.only}", file=sys.stderr, flush=True)
        if args.copy_probe_mb:
            copy_probe(args.copy_probe_mb)
        if args.only == "guiding":
            out = guiding_bench(ctx)
        elif args.only == "configs":
            out = configs_bench(ctx)
        else:
            out = traversal_bench(ctx, scenes.preset_scene("icosphere_renewal_512"))
        print(json.dumps(out), flush=True)
        ctx.close()
        return 0

    scene = scenes.preset_scene(PRESET)
    W, H, spp = scene["width"], scene["height"], scene["spp"]
    ctx = Context(0)
    s = Scene(ctx, scene)
    s.set_graph(not args.no_graph)  # the bounce loop replays a captured hipGraph
    print(f"bench: {ctx.info()} scene {W}x{H} spp {spp}", file=sys.stderr, flush=True)

    if args.copy_probe_mb:
        copy_probe(args.copy_probe_mb)

    for i in range(args.warmup):
        s.reset()
        s.render_passes(0, spp)
        print(f"bench: warmup {i} done", file=sys.stderr, flush=True)

    evals = 0
    steps_ref = 0
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        s.reset()
        s.render_passes(0, spp)
        st = s.stats()
        evals += st["field_evals"]
        steps_ref += st["march_steps"]
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    secs = t1 - t0
    # the march kernel's launches timed one by one (HIP events on the library
    # stream) on the same work, issued launch by launch: a replayed graph runs
    # the same kernels but is not timed per launch
    s.set_graph(False)
    march_ms, launches, evals_t = 0.0, 0, 0
    for i in range(min(args.steps, 3)):
        s.reset()
        s.render_passes(0, spp)
        st = s.stats()
        march_ms += st["march_kernel_ms"]
        launches += st["march_kernel_launches"]
        evals_t += st["field_evals"]
    paths = W * H * spp * args.steps
    ms_per_step = 1e3 * secs / args.steps
    ppc = scene["kernel"]["points_per_cell"]
    # a field evaluation: psi's value (16 flops per impulse) and the sphere mean (9);
    # the value-and-gradient sums of the medium hits the kernel shades (about one
    # per 140 evaluations on this workload) are not counted
    flops_per_eval = 27 * ppc * 16 + 9
    flops_t = evals_t * flops_per_eval
    achieved = flops_t / (march_ms * 1e-3) / 1e12 if march_ms > 0 else 0.0
    img = s.image()
    print(f"bench: {args.steps} steps in {secs:.3f} s; march kernel {march_ms:.1f} ms over "
          f"{launches} launches; {evals} field evals (reference grid points {steps_ref}); "
          f"image mean {img.mean():.5f}", file=sys.stderr, flush=True)

    guiding = None
    if not args.no_guiding:
        guiding = guiding_bench(ctx)
        print(f"bench: guiding {json.dumps(guiding)}", file=sys.stderr, flush=True)

    configs = None
    if not args.no_configs:
        configs = configs_bench(ctx)

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
                   "paths_per_step": W * H * spp, "tracer": "plain wavefront",
                   "bounce_loop": "launches issued one by one" if args.no_graph
                   else "hipGraph replay"},
        "roofline": {
            "bound": "valu",
            "kernel": f"k_march_sphere<{ppc}, 16>",
            "achieved": achieved,
            "peak": FP32_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / FP32_PEAK_TFLOPS,
            "traffic": MARCH_TRAFFIC_BYTES,
            "flops_per_launch": flops_t / max(1, launches),
            "field_evals_per_step": evals_t / min(args.steps, 3),
            # the reference's t-grid points (every one an evaluation in the reference's
            # march); the kernel evaluates the uncertain ones and decides the rest in
            # closed form, so the flops above count only the evaluated points
            "reference_grid_points_per_step": steps_ref / args.steps,
            "evaluated_fraction": (evals / steps_ref) if steps_ref else None,
            "avg_launch_ms": march_ms / max(1, launches),
        },
        "cpu_baseline": cpu,
        "guiding": guiding,
        "configs": configs,
    }
    print(json.dumps(line), flush=True)
    s.close()
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
