"""Benchmark: path-contributions/s of the NEE speckle tracer on one MI355X.

Workload (default `tm1024`): the 1024 x 1024 transmission-matrix render -- 128^3 voxel grid with
four Perlin-thresholded materials, HG g = 0.9, 1024 plane-wave illuminations x 1024 plane-wave
views, one frame, lambda = 1. A "step" is one batch of ``--batch`` paths (default 2^16) traced end
to end: free flight, NEE to every illumination and view, and the complex contraction into the
1024 x 1024 field. The timed region is ``--steps`` batches with the medium and tables already
resident in HBM; the value is paths * L * V * frames / wall time.

The measured form is the mixed-precision one: FP32 transport and terms, FP64 path length and
phase, the batch-relative attenuation cut (precision.tauCut, default 24: view-side terms
attenuated below exp(-12) of the batch's least attenuated view term are dropped) and the bf16x3
split-product MFMA contraction. The JSON line also carries:
  roofline              the dominant kernel (the view-term kernel) against HBM, with its VALU
                        issue rate beside it, from HIP events on its stream;
  roofline_contraction  the contraction's algorithmic MFMA rate in the pipeline, MFMA busy from the
                        counters, and its rate alone;
  accuracy              the field of the measured form against FP64 terms and FP64 contraction
                        without the cut, on one whole batch of the same paths (GPU), and against
                        the FP64 CPU restatement on a small batch (cpu_baseline leg);
  plain_gpu             the same tracer without split precision or cut (FP32 terms + FP32
                        contraction, and FP64 + FP64), with mixed/plain ratios;
  serving               a replay of the mixed request stream (three preloaded media, L and V in
                        [16, 512], 1-8 lags, Poisson arrivals) with p50/p99 latency;
  cpu_baseline          the FP64 CPU restatement (OpenMP) on a bounded sample of the workload.
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
# VALU issue peak for the unpacked instructions this library's kernels issue (packed FP32 is
# disabled): a wave64 instruction occupies a 16-lane SIMD for 4 cycles, so 256 CUs x 4 SIMDs x 16
# lanes x 2.4 GHz lane-instructions per second. MI355X's published 157.3 TFLOP/s FP32 vector
# figure counts packed FMAs (2 lanes' worth per lane, 2 flops each): 4x this in flops, 2x in FMAs.
# A microbenchmark of independent FP32 FMA chains issues 0.23 wave64 instructions per SIMD-cycle.
VALU_PEAK_LANE_INSTR = 256 * 4 * 16 * 2.4e9
METRIC = "path-contributions/sec (L×V complex) at err≤1e-3 vs FP64; p50/p99 render ms"


def workload(name: str, batch: int, steps: int):
    from temporalspecklescode_amd import workloads as W
    makers = {
        "tm1024": lambda: W.c4(paths=batch * steps, batch=batch),
        "hetero64": lambda: W.c2(paths=batch * steps, batch=batch),
        "slab2d": lambda: W.c1(paths=batch * steps, batch=batch),
        "corr128": lambda: W.c3(paths=batch * steps, batch=batch),
        "acorr128": lambda: W.c3_analytic(paths=batch * steps, batch=batch),
        # the confocal example's geometry: 1111 GaussianBeam illuminations, 3333 GaussianBeam views
        "confocal": lambda: W.confocal(batch=batch, full_iteration=steps),
    }
    if name not in makers:
        raise SystemExit("unknown workload %s" % name)
    return makers[name](), 2 if name == "slab2d" else 3


def host_cpu():
    """CPU model, the host's logical CPUs and the CPUs this process may use (its share of the box)."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count() or 1
    return model, os.cpu_count() or 1, usable


def cpu_baseline(name: str, seconds: float, threads: int, mixed_prec: dict):
    """The FP64 CPU restatement (OpenMP) rendering a growing sample of the workload's paths until
    it has run for about seconds/4; its rate is the baseline. The same leg checks the measured
    form's field against the restatement: a 128-path batch traced on the GPU in the measured form
    and the restatement's FP64 terms on the GPU's own vertices."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    from temporalspecklescode_amd import Tracer
    from temporalspecklescode_amd.config import sizes
    n = 8
    cfg, dims = workload(name, n, 1)
    nl, nv, nt, _ = sizes(cfg, dims)
    while True:
        cfg, dims = workload(name, n, 1)
        t0 = time.perf_counter()
        O.render(cfg, dims, nthreads=threads)
        elapsed = time.perf_counter() - t0
        if elapsed > seconds / 4 or n >= (1 << 20):
            break
        n *= 4
    rate = n / elapsed
    model, nproc, usable = host_cpu()
    out = {"value": rate * nl * nv * nt, "unit": "path-contributions/s", "cores": threads, "kind": "port",
           "sample": "%d paths of the same workload (the GPU run's seed, path ids 0..%d), %.1f s, FP64 scalar "
                     "restatement, OpenMP, %d threads" % (n, n - 1, elapsed, threads),
           "paths_per_s": rate, "cpu_model": model, "nproc": nproc, "cpus_usable": usable}
    # accuracy of the measured form against the restatement on the same vertices
    m = 128
    acfg, dims = workload(name, m, 1)
    acfg["precision"] = mixed_prec
    with Tracer(acfg, dims) as tr:
        d = tr.debug_batch(0, m)
    A, B, S = O.terms_from_vertices(dict(acfg, precision={"terms": "fp64", "contraction": "fp64"}), dims, d["k"],
                                    d["off"], d["pos"].astype(np.float64), d["mat"], d["rho"].astype(np.float64),
                                    d["xt"], d["len"], per_path_single=False)
    U = S.transpose(2, 1, 0).copy()
    for t in range(A.shape[1]):
        U[:, :, t] += A[:, t, :].T @ B[:, t, :]
    out["check_vs_oracle"] = {"paths": m, "rel_fro_err": float(np.linalg.norm(d["u"] - U) / np.linalg.norm(U))}
    return out


def accuracy(cfg: dict, dims: int, batch: int):
    """The measured form's batch field against FP64 terms + FP64 contraction without the cut, on
    the same batch of paths (path ids 0..batch-1)."""
    from temporalspecklescode_amd import Tracer
    a = dict(cfg)
    a["simulation"] = dict(cfg["simulation"], fullIteration=1, batchSize=batch)
    with Tracer(a, dims) as tm:
        dm = tm.debug_batch(0, batch)
    a["precision"] = {"terms": "fp64", "contraction": "fp64"}
    with Tracer(a, dims) as t64:
        d64 = t64.debug_batch(0, batch)
    err = float(np.linalg.norm(dm["u"] - d64["u"]) / max(1e-300, np.linalg.norm(d64["u"])))
    return {"rel_fro_err_vs_fp64": err, "paths": batch, "same_vertices": bool(np.array_equal(dm["k"], d64["k"])),
            "reference": "FP64 terms + FP64 contraction, no attenuation cut, same vertices (GPU; the FP64 path "
                         "matches the CPU restatement to 1e-10 at this geometry, tests/test_gpu_workloads.py)"}


def plain_gpu(cfg: dict, dims: int, batches: int, mixed_value: float):
    """The same tracer without split precision or cut: FP32 terms + FP32 contraction, and FP64
    terms + FP64 contraction, each timed over `batches` batches after one untimed trace."""
    from temporalspecklescode_amd import Tracer
    from temporalspecklescode_amd.config import sizes
    nl, nv, nt, _ = sizes(cfg, dims)
    batch = cfg["simulation"]["batchSize"]
    out = {}
    for name, prec in (("fp32", {"terms": "fp32", "contraction": "fp32"}),
                       ("fp64", {"terms": "fp64", "contraction": "fp64"})):
        c = dict(cfg, precision=prec)
        c["simulation"] = dict(cfg["simulation"], fullIteration=batches)
        with Tracer(c, dims) as tr:
            tr.trace()
            t0 = time.perf_counter()
            tr.trace()
            wall = time.perf_counter() - t0
        value = batch * batches * nl * nv * nt / wall
        out[name] = {"value": value, "ms_per_step": wall * 1e3 / batches, "steps": batches, "precision": prec}
        print("bench: plain %s %.3e path-contributions/s (%.1f ms/batch)" % (name, value, wall * 1e3 / batches),
              file=sys.stderr, flush=True)
    out["mixed_over_plain_fp32"] = mixed_value / out["fp32"]["value"]
    out["mixed_over_plain_fp64"] = mixed_value / out["fp64"]["value"]
    return out


def serving(n_requests: int, max_paths_log2: int, prec: dict):
    """Replays the mixed request stream (Poisson arrivals at 20/s, 2D/3D, three preloaded media,
    L and V log-uniform in [16, 512], 1-8 lags of linear motion) through the continuous-batching
    server in the measured form; latency percentiles and throughput."""
    from temporalspecklescode_amd import workloads as W
    from temporalspecklescode_amd.serving import Server
    stream = W.serving_stream(n_requests, max_paths_log2=max_paths_log2)
    for _, _, cfg in stream:
        cfg["precision"] = dict(prec)
    with Server(max_inflight=8) as srv:
        media = {}
        for _, dims, cfg in stream:   # preload each distinct medium once
            media.setdefault((dims, len(cfg["medium"]["xAxis"])), (dims, cfg))
        for dims, cfg in media.values():
            srv.preload(cfg, dims)
        t0 = time.perf_counter()
        ids = []
        for (arrival, dims, cfg) in stream:
            now = time.perf_counter() - t0
            if arrival > now:
                time.sleep(arrival - now)
            ids.append(srv.submit(cfg, dims))
        for rid in ids:
            srv.wait(rid)
            srv.fetch(rid)
        wall = time.perf_counter() - t0
        st = srv.stats()
    print("bench: served %d requests in %.2f s, p50 %.1f ms p99 %.1f ms" % (n_requests, wall, st["p50_ms"],
                                                                         st["p99_ms"]), file=sys.stderr, flush=True)
    return {"requests": n_requests, "p50_ms": st["p50_ms"], "p99_ms": st["p99_ms"], "mean_ms": st["mean_ms"],
            "paths_per_s": st["paths_per_s"], "contributions_per_s": st["contributions_per_s"], "wall_s": wall,
            "max_paths": 1 << max_paths_log2, "arrival_rate_per_s": 20.0, "preloaded_media": len(media),
            "precision": prec}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None, help="batches timed (default: the workload's full path "
                    "count for tm1024, 256 x 65536 = 2^24 paths; 8 for the others)")
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--batch", type=int, default=1 << 16)
    ap.add_argument("--workload", default="tm1024")
    ap.add_argument("--contraction", default="bf16x3")
    ap.add_argument("--tau-cut", type=float, default=24.0, help="precision.tauCut of the measured form (0: exact)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--no-accuracy", action="store_true")
    ap.add_argument("--serving", type=int, default=32, help="requests of the mixed stream to serve after the "
                    "throughput measurement (0: skip)")
    ap.add_argument("--serving-max-paths-log2", type=int, default=18)
    ap.add_argument("--plain-steps", type=int, default=2, help="batches timed for each plain (no split "
                    "precision, no cut) variant of the same workload (0: skip)")
    args = ap.parse_args()
    if args.steps is None:
        args.steps = (1 << 24) // args.batch if args.workload == "tm1024" else 8
    if args.gpus != 1:
        print("bench.py: this build runs and measures one GPU; paths would shard over GPUs by path-id range "
              "with one all-reduce of the field, which is not implemented, so --gpus %d is not measured"
              % args.gpus, file=sys.stderr)
        sys.exit(2)

    from temporalspecklescode_amd import Tracer
    from temporalspecklescode_amd.config import sizes

    cfg, dims = workload(args.workload, args.batch, args.steps)
    prec = {"terms": "fp32", "contraction": args.contraction, "tauCut": args.tau_cut}
    cfg["precision"] = prec
    nl, nv, nt, _ = sizes(cfg, dims)
    print("bench: workload %s, batch %d, steps %d, L %d V %d T %d, precision %s" % (
        args.workload, args.batch, args.steps, nl, nv, nt, json.dumps(prec)), file=sys.stderr, flush=True)
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

    # roofline of the dominant kernel: the multiple-scattering view terms (k_nee_view_far for
    # far-field views in FP32), timed by HIP events around its launches on its stream
    nb = max(1, tm["batches"])
    view_ms = tm["nee_view_ms"] / nb
    vtx_launch = tm["vertices"] / nb
    # algorithmic bytes of the tracing stage: 8 (L+V) T per path (complex64 term slabs written
    # once for the contraction) + (64 + 16 T) per vertex (vertex state in and out)
    alg_bytes = args.batch * 8.0 * (nl + nv) * nt + vtx_launch * (64.0 + 16.0 * nt)
    gbs = alg_bytes / (view_ms * 1e-3) / 1e9
    steps_s = tm["dda_steps_view"] / max(1e-9, tm["nee_view_ms"] * 1e-3)
    prof = {}
    tpath = os.path.join(ROOT, "profiles", "counters_%s.json" % args.workload)
    if os.path.exists(tpath):
        with open(tpath) as f:
            prof = json.load(f)
    same_form = prof.get("batch") == args.batch and prof.get("precision") == prec
    roof = {"kernel": prof.get("view_kernel", "k_nee_view_far"), "bound": "hbm", "achieved": round(gbs, 3),
            "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 6),
            "traffic": prof.get("view_hbm_bytes") if same_form else None, "avg_launch_ms": round(view_ms, 3),
            "algorithmic_bytes_per_launch": alg_bytes, "voxel_steps_per_s": steps_s,
            "voxel_steps_per_launch": tm["dda_steps_view"] / nb,
            "first_vertex_voxel_steps_per_batch": (tm["dda_steps"] - tm["dda_steps_view"]) / nb,
            "limiter": "valu issue (the walk's grid reads hit L2; see valu)"}
    # the walk is bound by VALU issue, not HBM: VALU lane-instructions per voxel step (SQ_INSTS_VALU
    # over the same kernel, inactive lanes included) against the unpacked VALU issue peak
    if same_form and "view_valu_lane_instr_per_step" in prof:
        ips = prof["view_valu_lane_instr_per_step"]
        ach = ips * steps_s
        roof["valu"] = {"lane_instr_per_step": ips, "achieved": ach, "peak": VALU_PEAK_LANE_INSTR,
                        "unit": "lane-instr/s", "frac": round(ach / VALU_PEAK_LANE_INSTR, 4)}
    flops = 8.0 * args.batch * nl * nv * nt   # per contraction launch set: 8 flops per path-contribution
    con_ms = tm["contract_ms"] / nb
    con_tflops = flops / (con_ms * 1e-3) / 1e12
    # the multiple-scattering contraction as it runs in the pipeline (split, GEMM and reduce, timed
    # by events on the contraction stream while the next batch's NEE runs on another stream)
    roof_con = {"kernel": "k_row_nonzero + k_compact_paths + k_split_bf16 + k_contract_bf16x3_planes + k_reduce"
                if args.contraction == "bf16x3"
                else "k_contract_fma",
                "bound": "mfma", "achieved": round(con_tflops, 3), "peak": MFMA_BF16_PEAK_TFLOPS,
                "unit": "TFLOP/s", "frac": round(con_tflops / MFMA_BF16_PEAK_TFLOPS, 6), "traffic": None,
                "avg_launch_ms": round(con_ms, 4), "timed": "in the pipeline, concurrent with the NEE stage"}
    if same_form and "contract_mfma_busy_pct" in prof:
        roof_con["mfma_busy_pct_counters"] = prof["contract_mfma_busy_pct"]
        # HBM bytes of one contraction: the GEMM's plus both operands' splits
        roof_con["traffic"] = (prof.get("contract_fetch_bytes", 0) + prof.get("contract_write_bytes", 0) +
                               2 * (prof.get("split_fetch_bytes", 0) + prof.get("split_write_bytes", 0)))
    if same_form and "contract_alone_total_ms" in prof:
        roof_con["alone"] = {"ms": prof["contract_alone_total_ms"], "tflops": prof["contract_alone_tflops"],
                             "gemm_tflops": prof["contract_alone_gemm_tflops"],
                             "source": "profiles/counters_tm1024.json (kernel trace, no concurrent work)"}

    out = {
        "metric": METRIC,
        "value": value, "unit": "path-contributions/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp32 transport+terms, fp64 phase, bf16x3 contraction" if args.contraction == "bf16x3" else
                 "fp32 transport+terms, fp64 phase, %s contraction" % args.contraction,
        "data": "synthetic (seeded Perlin medium, counter-based Philox paths)",
        "config": {"workload": args.workload, "batch_paths": args.batch, "L": nl, "V": nv, "frames": nt,
                   "paths": paths, "contraction": args.contraction, "tau_cut": args.tau_cut},
        "paths_per_s": paths / wall, "vertices_per_s": tm["vertices"] / wall,
        "roofline": roof, "roofline_contraction": roof_con,
        "stage_ms": {k: tm[k] for k in ("transport_ms", "nee_ms", "nee_view_ms", "contract_ms", "single_ms",
                                        "total_ms")},
    }
    if not args.no_accuracy:
        out["accuracy"] = accuracy(cfg, dims, args.batch)
        print("bench: accuracy %s" % json.dumps(out["accuracy"]), file=sys.stderr, flush=True)
    if args.plain_steps > 0:
        out["plain_gpu"] = plain_gpu(cfg, dims, args.plain_steps, value)
    if args.serving > 0:
        out["serving"] = serving(args.serving, args.serving_max_paths_log2, prec)
    if not args.no_cpu:
        _, _, usable = host_cpu()
        threads = int(os.environ.get("OMP_NUM_THREADS", usable))
        threads = max(1, min(threads, usable))
        out["cpu_baseline"] = cpu_baseline(args.workload, args.cpu_seconds, threads, prec)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
