"""Benchmark: path-contributions/s of the NEE speckle tracer on one MI355X.

Workload (default): the 1024 x 1024 transmission-matrix render -- 128^3 voxel grid with four
Perlin-thresholded materials, HG g = 0.9, 1024 plane-wave illuminations x 1024 plane-wave views,
one frame, lambda = 1. A "step" is one batch of ``--batch`` paths (default 2^16) traced end to end:
free flight, NEE to every illumination and view, and the complex contraction into the 1024 x 1024
field. The timed region is ``--steps`` batches on the device with the medium and tables already
resident in HBM; the value is paths * L * V * frames / wall time.

The JSON line also carries:
  roofline      the dominant kernel's achieved rate against its MI355X peak (HIP events on the
                tracer's stream); the contraction's MFMA rate is reported beside it;
  accuracy      relative Frobenius error of the mixed-precision field against the FP64 terms +
                FP64 contraction on the same paths (one batch);
  cpu_baseline  the FP64 CPU restatement (OpenMP) on a bounded sample of the same workload;
  plain_gpu     the same tracer without split precision (FP32 terms + FP32 contraction, and FP64
                terms + FP64 contraction) on the same workload, with mixed/plain ratios.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
MFMA_BF16_PEAK_TFLOPS = 2500.0
VALU_PEAK_LANE_INSTR = 157.3e12 / 2       # published FP32 vector peak in lane-FMAs (2 flops each) per s


def workload(name: str, batch: int, steps: int):
    from temporalspecklescode_amd import workloads as W
    if name == "tm1024":
        cfg = W.c4(paths=batch * steps, batch=batch)
        return cfg, 3
    if name == "hetero64":
        cfg = W.c2(paths=batch * steps, batch=batch)
        return cfg, 3
    if name == "slab2d":
        cfg = W.c1(paths=batch * steps, batch=batch)
        return cfg, 2
    if name == "corr128":
        cfg = W.c3(paths=batch * steps, batch=batch)
        return cfg, 3
    if name == "acorr128":
        cfg = W.c3_analytic(paths=batch * steps, batch=batch)
        return cfg, 3
    if name == "confocal":
        # the confocal example's geometry: 1111 GaussianBeam illuminations, 3333 GaussianBeam views
        cfg = W.confocal(batch=batch, full_iteration=steps)
        return cfg, 3
    raise SystemExit("unknown workload %s" % name)


def cpu_baseline(name: str, seconds: float, threads: int):
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    from temporalspecklescode_amd.config import sizes
    n = 8
    cfg, dims = workload(name, n, 1)
    nl, nv, nt, _ = sizes(cfg, dims)
    elapsed = 0.0
    while True:
        cfg, dims = workload(name, n, 1)
        t0 = time.perf_counter()
        O.render(cfg, dims, nthreads=threads)
        elapsed = time.perf_counter() - t0
        if elapsed > seconds / 4 or n >= (1 << 20):
            break
        n *= 4
    rate = n / elapsed
    return {"value": rate * nl * nv * nt, "unit": "path-contributions/s", "cores": threads, "kind": "port",
            "sample": "%d paths of the same workload (seed and batch as the GPU run's first paths), %.1f s, "
                      "FP64 scalar restatement, OpenMP" % (n, elapsed),
            "paths_per_s": rate}


def plain_gpu(cfg: dict, dims: int, batches: int, mixed_value: float):
    """The same tracer without split precision (the plain GPU tracer): FP32 transport, FP64
    path length and phase, and (i) FP32 terms with an FP32 FMA contraction, (ii) FP64 terms with
    an FP64 contraction. Each is timed over `batches` batches (the kernels are already loaded by the
    mixed run) and its field error is taken against FP64 terms + FP64 contraction on the same
    4096 paths."""
    from temporalspecklescode_amd import Tracer
    from temporalspecklescode_amd.config import sizes
    nl, nv, nt, _ = sizes(cfg, dims)
    batch = cfg["simulation"]["batchSize"]
    out = {}
    ref = None
    for name, prec in (("fp64", {"terms": "fp64", "contraction": "fp64"}), ("fp32", {"terms": "fp32", "contraction": "fp32"})):
        c = dict(cfg, precision=prec)
        c["simulation"] = dict(cfg["simulation"], fullIteration=batches)
        with Tracer(c, dims) as tr:
            t0 = time.perf_counter()
            tr.trace()
            wall = time.perf_counter() - t0
        a = dict(c)
        n = min(batch, 4096)
        a["simulation"] = dict(cfg["simulation"], fullIteration=1, batchSize=n)
        with Tracer(a, dims) as td:
            u = td.debug_batch(0, n)["u"]
        if ref is None:
            ref = u
        err = float(np.linalg.norm(u - ref) / max(1e-300, np.linalg.norm(ref)))
        value = batch * batches * nl * nv * nt / wall
        out[name] = {"value": value, "ms_per_step": wall * 1e3 / batches, "steps": batches,
                     "rel_fro_err_vs_fp64": err, "precision": prec}
        print("bench: plain %s %.3e path-contributions/s (%.1f ms/batch)" % (name, value, wall * 1e3 / batches),
              file=sys.stderr, flush=True)
    out["mixed_over_plain_fp32"] = mixed_value / out["fp32"]["value"]
    out["mixed_over_plain_fp64"] = mixed_value / out["fp64"]["value"]
    return out


def serving(n_requests: int, max_paths_log2: int):
    """Replays the mixed request stream (Poisson arrivals, 2D/3D, three media, L and V in
    [16, 512], 1-8 lags of linear motion) through the continuous-batching server; returns
    latency percentiles and throughput."""
    from temporalspecklescode_amd import workloads as W
    from temporalspecklescode_amd.serving import Server
    from temporalspecklescode_amd.config import sizes
    stream = W.serving_stream(n_requests, max_paths_log2=max_paths_log2)
    t0 = time.perf_counter()
    ids = []
    contrib = 0
    with Server(max_inflight=8) as srv:
        for (arrival, dims, cfg) in stream:
            now = time.perf_counter() - t0
            if arrival > now:
                time.sleep(arrival - now)
            ids.append(srv.submit(cfg, dims))
            nl, nv, nt, _ = sizes(cfg, dims)
            contrib += cfg["simulation"]["batchSize"] * cfg["simulation"]["fullIteration"] * nl * nv * nt
        for rid in ids:
            srv.wait(rid)
            srv.fetch(rid)
        wall = time.perf_counter() - t0
        st = srv.stats()
    print("bench: served %d requests in %.2f s, p50 %.1f ms p99 %.1f ms" % (n_requests, wall, st["p50_ms"],
                                                                         st["p99_ms"]), file=sys.stderr, flush=True)
    return {"requests": n_requests, "p50_ms": st["p50_ms"], "p99_ms": st["p99_ms"], "mean_ms": st["mean_ms"],
            "paths_per_s": st["paths_per_s"], "contributions_per_s": st["contributions_per_s"],
            "wall_s": wall, "max_paths": 1 << max_paths_log2, "arrival_rate_per_s": 20.0}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=8)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--batch", type=int, default=1 << 16)
    ap.add_argument("--workload", default="tm1024")
    ap.add_argument("--contraction", default="bf16x3")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--no-accuracy", action="store_true")
    ap.add_argument("--serving", type=int, default=0, help="requests of the mixed stream to serve after the "
                    "throughput measurement (0: skip)")
    ap.add_argument("--serving-max-paths-log2", type=int, default=18)
    ap.add_argument("--plain-steps", type=int, default=2, help="batches timed for each plain (no split "
                    "precision) variant of the same workload (0: skip)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: runs on one GPU; sharding paths over GPUs by path-id range is not implemented, "
              "so --gpus %d is not measured" % args.gpus, file=sys.stderr)
        sys.exit(2)

    from temporalspecklescode_amd import Tracer
    from temporalspecklescode_amd.config import sizes

    cfg, dims = workload(args.workload, args.batch, args.steps)
    cfg["precision"] = {"terms": "fp32", "contraction": args.contraction}
    nl, nv, nt, _ = sizes(cfg, dims)
    print("bench: workload %s, batch %d, steps %d, L %d V %d T %d" % (args.workload, args.batch, args.steps,
                                                                     nl, nv, nt), file=sys.stderr, flush=True)
    if args.warmup > 0:
        wcfg = dict(cfg)
        wcfg["simulation"] = dict(cfg["simulation"], fullIteration=args.warmup, seed=cfg["simulation"]["seed"] + 1000)
        with Tracer(wcfg, dims) as tw:
            tw.trace()
        print("bench: warmup done", file=sys.stderr, flush=True)
    with Tracer(cfg, dims) as tr:
        t0 = time.perf_counter()
        u, it = tr.trace()          # synchronous: returns after the device finished
        wall = time.perf_counter() - t0
        tm = tr.timing()
    paths = args.batch * args.steps
    contributions = paths * nl * nv * nt
    value = contributions / wall
    print("bench: timed %.3f s, stages %s" % (wall, json.dumps(tm)), file=sys.stderr, flush=True)

    # roofline of the dominant kernel (k_nee_view, the multiple-scattering view terms: >90% of the
    # device time), timed by HIP events around its launches on the tracer's stream
    nb = max(1, tm["batches"])
    view_ms = tm["nee_view_ms"] / nb
    con_ms = (tm["contract_ms"] + tm["single_ms"]) / nb
    vtx_launch = tm["vertices"] / nb
    # algorithmic bytes of the tracing stage: 8 (L+V) T per path (complex64 term slabs written
    # once for the contraction) + (64 + 16 T) per vertex (vertex state in and out)
    alg_bytes = args.batch * 8.0 * (nl + nv) * nt + vtx_launch * (64.0 + 16.0 * nt)
    gbs = alg_bytes / (view_ms * 1e-3) / 1e9
    steps_s = tm["dda_steps_view"] / max(1e-9, tm["nee_view_ms"] * 1e-3)   # k_nee_view's own walks
    prof = {}
    tpath = os.path.join(ROOT, "profiles", "counters_%s.json" % args.workload)
    if os.path.exists(tpath):
        with open(tpath) as f:
            prof = json.load(f)
    traffic = None
    if prof.get("batch") == args.batch and "k_nee_view_hbm_bytes" in prof:
        traffic = prof["k_nee_view_hbm_bytes"]
    roof = {"kernel": "k_nee_view", "bound": "hbm", "achieved": round(gbs, 3), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(gbs / HBM_PEAK_GBS, 6), "traffic": traffic, "avg_launch_ms": round(view_ms, 3),
            "algorithmic_bytes_per_launch": alg_bytes, "voxel_steps_per_s": steps_s}
    # the walk is VALU-issue-bound, not HBM-bound: VALU lane-instructions per voxel step (from the
    # SQ_INSTS_VALU counter over the same kernel, inactive lanes included) against MI355X's
    # published FP32 vector peak counted in lane-instructions (157.3 TFLOP/s / 2 flops per FMA)
    if "k_nee_view_valu_lane_instr_per_step" in prof:
        ips = prof["k_nee_view_valu_lane_instr_per_step"]
        ach = ips * steps_s
        roof["valu"] = {"lane_instr_per_step": ips, "achieved": ach, "peak": VALU_PEAK_LANE_INSTR,
                        "unit": "lane-instr/s", "frac": round(ach / VALU_PEAK_LANE_INSTR, 4)}
    flops = 2 * 8.0 * args.batch * nl * nv * nt   # multiple + single contraction, 8 flops each
    con_tflops = flops / (con_ms * 1e-3) / 1e12
    roof_con = {"kernel": "k_contract_bf16x3" if args.contraction == "bf16x3" else "k_contract_fma",
                "bound": "mfma", "achieved": round(con_tflops, 3), "peak": MFMA_BF16_PEAK_TFLOPS,
                "unit": "TFLOP/s", "frac": round(con_tflops / MFMA_BF16_PEAK_TFLOPS, 6), "traffic": None,
                "avg_launch_ms": round(con_ms / 2, 4)}

    accuracy = None
    if not args.no_accuracy:
        acfg = dict(cfg)
        acfg["simulation"] = dict(cfg["simulation"], fullIteration=1)
        n = min(args.batch, 4096)
        acfg["simulation"]["batchSize"] = n
        with Tracer(acfg, dims) as tm32:
            d32 = tm32.debug_batch(0, n)
        acfg["precision"] = {"terms": "fp64", "contraction": "fp64"}
        with Tracer(acfg, dims) as tm64:
            d64 = tm64.debug_batch(0, n)
        same = np.array_equal(d32["k"], d64["k"])
        err = float(np.linalg.norm(d32["u"] - d64["u"]) / max(1e-300, np.linalg.norm(d64["u"])))
        accuracy = {"rel_fro_err_vs_fp64": err, "paths": n, "same_vertices": bool(same)}

    out = {
        "metric": "path-contributions/sec (L×V complex) at err≤1e-3 vs FP64; p50/p99 render ms",
        "value": value, "unit": "path-contributions/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp32 transport+terms, fp64 phase, bf16x3 contraction" if args.contraction == "bf16x3" else "fp32",
        "data": "synthetic (seeded Perlin medium, counter-based Philox paths)",
        "config": {"workload": args.workload, "batch_paths": args.batch, "L": nl, "V": nv, "frames": nt,
                   "paths": paths, "contraction": args.contraction},
        "paths_per_s": paths / wall, "vertices_per_s": tm["vertices"] / wall,
        "roofline": roof, "roofline_contraction": roof_con,
        "accuracy": accuracy,
        "stage_ms": {k: tm[k] for k in ("transport_ms", "nee_ms", "nee_view_ms", "contract_ms", "single_ms")},
    }
    if args.plain_steps > 0:
        out["plain_gpu"] = plain_gpu(cfg, dims, args.plain_steps, value)
    if args.serving > 0:
        out["serving"] = serving(args.serving, args.serving_max_paths_log2)
    if not args.no_cpu:
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        threads = max(1, min(threads, 16))
        out["cpu_baseline"] = cpu_baseline(args.workload, args.cpu_seconds, threads)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
