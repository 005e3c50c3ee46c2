#!/usr/bin/env python3
"""Benchmark: 1M x 128 k-NN graph build points/s (BF16-MFMA join vs FP32 join) at recall@15,
plus batched query QPS and p50/p99 latency, on one MI355X.

Workload: 1,000,000 x 128 float32 integer-valued Gaussian mixture (1,000 centres U[0,100)^128,
sigma 5, rounded and clipped to 0..255), squared Euclidean, k=15, RP forest (16 trees, leaf 256) +
NN-Descent (max_candidates 15, <= 20 iterations, delta 0.001), counter-based draws (the GPU-native
sampling mode).  A step is one full build from HBM-resident data to the sorted k-NN graph (forest,
leaf init, iterations, sort_all).  One extra FP32 build with the reference's FastRng streams (the
bit-exact mode) is timed and reported beside it.

Prints one JSON line.  `value` is the points/s of the build with the headline join precision.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=1_000_000)
    ap.add_argument("--dim", type=int, default=128)
    ap.add_argument("--k", type=int, default=15)
    ap.add_argument("--trees", type=int, default=16)
    ap.add_argument("--leaf", type=int, default=256)
    ap.add_argument("--iters", type=int, default=20)
    ap.add_argument("--delta", type=float, default=0.001)
    ap.add_argument("--threads", type=int, default=8, help="Rayon thread count reproduced by candidate sampling")
    ap.add_argument("--rng", default="counter", choices=["reference", "counter"])
    ap.add_argument("--no-ref", action="store_true", help="skip the reference-stream FP32 build")
    ap.add_argument("--no-defaults", action="store_true",
                    help="skip the builds at the reference's default n_trees / leaf_size")
    ap.add_argument("--precision", default="both", choices=["fp32", "bf16", "both"])
    ap.add_argument("--recall-rows", type=int, default=10_000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-query", action="store_true")
    ap.add_argument("--queries", type=int, default=10_000)
    ap.add_argument("--profile", action="store_true", help="builds only (for rocprofv3 runs)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="first run a plain 1 GiB device-to-device copy (3 times): the known-size kernel that "
                         "scales FETCH_SIZE / WRITE_SIZE counters")
    ap.add_argument("--epsilons", default="0.1,0.15,0.2",
                    help="comma-separated query epsilons; the fastest run at recall@10 >= 0.95 is reported")
    ap.add_argument("--fp8-epsilons", default="0.1,0.15",
                    help="epsilons of the FP8-candidate search with exact re-ranking (empty: none)")
    ap.add_argument("--no-precision-form", action="store_true",
                    help="skip the 1,000,000 x 960, k=30 precision-form builds (BF16 join vs FP32)")
    ap.add_argument("--pf-n", type=int, default=1_000_000, help="rows of the precision-form workload")
    ap.add_argument("--pf-steps", type=int, default=2, help="timed builds per precision of the precision form")
    ap.add_argument("--leg", choices=["batching", "precision", "serving"],
                    help="run only this form's leg and print its JSON (profiling runs)")
    ap.add_argument("--no-batching", action="store_true", help="skip the batching-form leg (512 independent indexes)")
    ap.add_argument("--bf-count", type=int, default=512, help="indexes of the batching form")
    ap.add_argument("--bf-steps", type=int, default=2, help="timed batched builds of the batching form")
    ap.add_argument("--bf-no-sequential", action="store_true",
                    help="skip the sequential builds of the batching form (profiling runs)")
    ap.add_argument("--no-serving", action="store_true",
                    help="skip the serving form (2,000,000 x 96 resident index, Poisson query stream)")
    ap.add_argument("--sv-n", type=int, default=2_000_000, help="rows of the serving-form index")
    ap.add_argument("--sv-queries", type=int, default=1_000_000,
                    help="queries of the serving-form stream (the stated form has 4,000,000; the default run streams "
                         "1,000,000 of the same form to stay within a few minutes)")
    ap.add_argument("--sv-load", type=float, default=0.7, help="offered Poisson rate as a fraction of the saturated QPS")
    ap.add_argument("--sv-modes", default="fp32,fp32_w32,fp8,int8",
                    help="searches of the serving form: fp32 (the reference's query), fp32_wN (search-width floor N), "
                         "fp8 (FP8 candidates + FP32 re-ranking), int8 (the reference's quantized cosine index: int8 "
                         "codes, default widths)")
    ap.add_argument("--sv-visited-gb", type=float, default=8.0,
                    help="visited-set pool of the query service (GiB; n bytes per concurrent search)")
    ap.add_argument("--sv-inflight", type=int, default=4, help="launches in flight in the query service")
    ap.add_argument("--sv-target", type=int, default=1024, help="queries at which a launch of the query service closes")
    return ap.parse_args()


def gpu_build(ctx, L, X, n, dim, args, precision, outs):
    idx, dist, flags = outs
    p = L.make_params(args.k, args.trees, args.leaf, min(args.k, 60), args.iters, args.delta, n_threads=args.threads,
                      rng_mode=args.rng, join_precision=precision)
    st = L.rng_seed(42)
    stats = L.BuildStats()
    h = C.c_void_p()
    rc = L.lib().rinnd_nn_descent(ctx.h, L.ptr(X), n, dim, L.DIST["sqeuclidean"], C.byref(p), L.ptr(st), L.ptr(idx),
                                  L.ptr(dist), L.ptr(flags), C.byref(h), C.byref(stats))
    ctx.check(rc, "rinnd_nn_descent")
    L.lib().rinnd_forest_free(h)
    ctx.check(L.lib().rinnd_sort_rows(ctx.h, L.ptr(idx), L.ptr(dist), L.ptr(flags), n, args.k), "rinnd_sort_rows")
    return stats.to_dict()


def exact_knn_f64(torch, Xd, rows, k, chunk=1024):
    """Exact k nearest neighbours (excluding self) of Xd[rows] in float64 on the GPU."""
    X64 = Xd.double()
    sq = (X64 * X64).sum(1)
    out_d, out_i = [], []
    for s in range(0, len(rows), chunk):
        r = torch.as_tensor(rows[s:s + chunk], device=Xd.device)
        q = X64[r]
        D = sq[r][:, None] + sq[None, :] - 2.0 * (q @ X64.T)
        D[torch.arange(len(r), device=Xd.device), r] = float("inf")
        d, i = torch.topk(D, k, dim=1, largest=False)
        out_d.append(d)
        out_i.append(i)
    return torch.cat(out_d), torch.cat(out_i)


def exact_knn_queries_f64(torch, Xd, Q, k, chunk=1024):
    X64 = Xd.double()
    sq = (X64 * X64).sum(1)
    out_d, out_i = [], []
    for s in range(0, Q.shape[0], chunk):
        q = Q[s:s + chunk].double()
        D = (q * q).sum(1)[:, None] + sq[None, :] - 2.0 * (q @ X64.T)
        d, i = torch.topk(D, k, dim=1, largest=False)
        out_d.append(d)
        out_i.append(i)
    return torch.cat(out_d), torch.cat(out_i)


def tie_aware_recall(torch, Xd, src, found_idx, true_d, k):
    """hit if the FP64 distance of a returned id is <= the k-th true FP64 distance."""
    X64 = Xd.double()
    f = found_idx.long().clamp(min=0)
    d = ((X64[f] - src.double()[:, None, :]) ** 2).sum(-1)
    d[found_idx < 0] = float("inf")
    hits = (d <= true_d[:, k - 1:k] * (1 + 1e-12)).sum(1).double() / k
    return float(hits.mean())


def cpu_baseline(X, args):
    """The oracle's nn_descent + sort_all (the GPU timing's span) on a bounded sample, on every host core this job
    may use: the cores of its affinity mask, capped at the host-CPU share of one GPU on the measurement machine
    (16 threads)."""
    from oracle import oracle as O
    visible = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    cores = max(1, min(visible, 16))
    sample = min(args.n, 50_000)
    t = 0.0
    while True:
        Xs = np.ascontiguousarray(X[:sample])
        t0 = time.perf_counter()
        r = O.nn_descent(Xs, args.k, args.trees, args.leaf, min(args.k, 60), args.iters, 0.001, threads=cores)
        O.sort_all(r["indices"], r["distances"], r["flags"])
        t = time.perf_counter() - t0
        if t > 8.0 or sample >= args.n:
            break
        sample = min(args.n, sample * 2 if t > 2.0 else sample * 4)
    return {"value": sample / t, "unit": "points/s", "cores": cores, "kind": "port",
            "sample": f"oracle nn_descent + sort_all on the first {sample} of {args.n} rows of the same data, same "
                      f"parameters, {cores} threads of {visible} visible ({t:.2f} s)"}


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures the single-GPU build only (--gpus {args.gpus} requested): the k-NN build "
              "runs on one GPU, there is no sharded multi-GPU path", file=sys.stderr)
        sys.exit(2)
    import torch
    from rinnd_amd import _lib as L
    from rinnd_amd import datasets

    if args.leg:
        fn = {"batching": batching_form, "precision": precision_form, "serving": serving_form}[args.leg]
        print(json.dumps({args.leg + "_form": fn(torch, L.Context.default(0), L, args)}), flush=True)
        return
    n, dim, k = args.n, args.dim, args.k
    if args.copy_calibration:
        src = torch.ones(1 << 28, dtype=torch.float32, device="cuda")  # 1 GiB
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    log(f"[bench] generating {n}x{dim} integer Gaussian mixture")
    X = datasets.cfg2(n=n, d=dim)
    Xd = torch.from_numpy(X).cuda()
    ctx = L.Context.default(0)
    outs = (torch.empty((n, k), dtype=torch.int32, device="cuda"),
            torch.empty((n, k), dtype=torch.float32, device="cuda"),
            torch.empty((n, k), dtype=torch.uint8, device="cuda"))
    precisions = ["fp32", "bf16"] if args.precision == "both" else [args.precision]
    results = {}
    for prec in precisions:
        try:
            for w in range(args.warmup):
                log(f"[bench] warmup {prec} {w + 1}/{args.warmup}")
                gpu_build(ctx, L, Xd, n, dim, args, prec, outs)
            torch.cuda.synchronize()
            times, stats = [], []
            for s in range(args.steps):
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                st = gpu_build(ctx, L, Xd, n, dim, args, prec, outs)
                torch.cuda.synchronize()
                times.append(time.perf_counter() - t0)
                stats.append(st)
                log(f"[bench] {prec} step {s + 1}: {times[-1] * 1e3:.1f} ms, iters {st['iterations']}, "
                    f"forest {st['ms_forest']:.1f} init {st['ms_init']:.1f} cand {st['ms_candidates']:.1f} "
                    f"join {st['ms_join']:.1f} apply {st['ms_apply']:.1f} ms")
                it = st["iterations"]
                log(f"[bench]   per iteration: pairs {st['join_pairs'][:it]} updates {st['join_updates'][:it]} "
                    f"changes {st['changes'][:it]} rescored {st['join_rescored'][:it]}")
            results[prec] = {"times": times, "stats": stats, "idx": outs[0].clone()}
        except Exception as e:  # report what ran
            log(f"[bench] {prec} build failed: {e}")
    if not results:
        raise SystemExit("no build succeeded")

    rows = np.random.default_rng(3).choice(n, size=min(args.recall_rows, n), replace=False)
    recall = {}
    true_d = None
    if not args.profile:
        log("[bench] exact FP64 k-NN of the recall rows")
        true_d, true_i = exact_knn_f64(torch, Xd, rows, k)
        rt = torch.as_tensor(rows, device="cuda")
        for prec, r in results.items():
            recall[prec] = tie_aware_recall(torch, Xd, Xd[rt], r["idx"][rt], true_d, k)
            log(f"[bench] recall@{k} {prec}: {recall[prec]:.4f}")

    head = "bf16" if "bf16" in results else "fp32"
    T = results[head]["times"]
    ms = 1e3 * sum(T) / len(T)
    st = results[head]["stats"]
    roofline = roofline_of(st, head, dim)
    kernels = {fam: kernel_line(st, fam, head, dim) for fam in L.TIMED_KERNELS}

    line = {
        "metric": "1Mx128 kNN build pts/s BF16-join vs FP32 at recall@15; query QPS + p99",
        "value": n / (ms / 1e3), "unit": "points/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32" if head == "fp32" else "bf16-join/f32", "data": "synthetic",
        "config": {"workload": f"{n}x{dim} integer Gaussian mixture (1000 centres, sigma 5), sqeuclidean, "
                               f"k={k}, {args.trees} trees, leaf {args.leaf}, <= {args.iters} iterations",
                   "head_precision": head, "rng_mode": args.rng, "n_threads": args.threads},
        "builds": {p: {"points_per_s": n / (sum(r["times"]) / len(r["times"])),
                       "ms_per_build": 1e3 * sum(r["times"]) / len(r["times"]),
                       "recall_at_k": recall.get(p), "iterations": r["stats"][-1]["iterations"],
                       "phase_ms": {x: r["stats"][-1][x] for x in
                                    ("ms_forest", "ms_init", "ms_candidates", "ms_join", "ms_apply")}}
                   for p, r in results.items()},
        "roofline": roofline,
        "kernels": kernels,
    }
    if "fp32" in results and "bf16" in results:
        line["bf16_speedup"] = line["builds"]["bf16"]["points_per_s"] / line["builds"]["fp32"]["points_per_s"]

    if not args.profile and not args.no_defaults:
        try:
            line["reference_defaults"] = defaults_build(torch, ctx, L, Xd, n, dim, args, outs, rows, true_d)
        except Exception as e:
            log(f"[bench] reference-defaults builds failed: {e}")
    if not args.profile and not args.no_ref and args.rng != "reference":
        try:
            line["reference_stream_build"] = reference_build(torch, ctx, L, Xd, n, dim, args, outs, rows, true_d)
        except Exception as e:
            log(f"[bench] reference-stream build failed: {e}")
    if not args.profile and not args.no_query:
        try:
            line["query"] = query_bench(torch, L, ctx, X, Xd, args)
        except Exception as e:
            log(f"[bench] query bench failed: {e}")
    if not args.profile and not args.no_precision_form:
        try:
            line["precision_form"] = precision_form(torch, ctx, L, args)
        except Exception as e:
            log(f"[bench] precision form failed: {e}")
    if not args.profile and not args.no_batching:
        try:
            line["batching_form"] = batching_form(torch, ctx, L, args)
        except Exception as e:
            log(f"[bench] batching form failed: {e}")
    if not args.profile and not args.no_serving:
        try:
            line["serving_form"] = serving_form(torch, ctx, L, args)
        except Exception as e:
            log(f"[bench] serving form failed: {e}")
    if not args.profile and not args.no_cpu:
        try:
            line["cpu_baseline"] = cpu_baseline(X, args)
        except Exception as e:
            log(f"[bench] cpu baseline failed: {e}")
    print(json.dumps(line), flush=True)


# timed family -> kernel (per join precision; the synthetic data are integers, exact in bf16, so the BF16
# build's leaf distances run on MFMA)
KERNEL_NAMES = {"join": {"fp32": "k_join_fp32", "bf16": "k_join_bf16"}, "margins": "k_lvl_margins",
                "leaf_dist": {"fp32": "k_leaf_dist_sq", "bf16": "k_leaf_dist_mfma/k_leaf_surv_mfma"},
                "leaf_push": "k_leaf_push"}


MFMA_BF16_PEAK_TFLOPS = 2500.0  # dense
FP32_VECTOR_PEAK_TFLOPS = 157.3


def kernel_line(stats, fam, head, dim):
    ms = sum(s["timed"][fam]["ms"] for s in stats)
    launches = sum(s["timed"][fam]["launches"] for s in stats)
    nbytes = sum(s["timed"][fam]["bytes"] for s in stats)
    name = KERNEL_NAMES[fam][head] if isinstance(KERNEL_NAMES[fam], dict) else KERNEL_NAMES[fam]
    if not launches or ms <= 0:
        return {"kernel": name, "launches": 0}
    achieved = (nbytes / launches) / (ms / launches * 1e-3) / 1e9
    out = {"kernel": name, "launches": launches // len(stats), "ms_per_build": ms / len(stats),
           "avg_launch_ms": ms / launches, "bytes_per_launch": nbytes / launches, "achieved_gbs": achieved,
           "frac_hbm": achieved / HBM_PEAK_GBS}
    if fam == "join":  # algorithmic flops of the join tiles: 2 d per candidate pair (the MFMA or FP32 roof)
        flops = sum(2.0 * dim * sum(s["join_pairs"][:s["iterations"]]) for s in stats)
        tflops = flops / (ms * 1e-3) / 1e12
        peak = MFMA_BF16_PEAK_TFLOPS if head == "bf16" else FP32_VECTOR_PEAK_TFLOPS
        out.update({"flops_per_launch": flops / launches, "achieved_tflops": tflops,
                    "peak_tflops": peak, "frac_compute_peak": tflops / peak,
                    "compute_peak": "dense BF16 MFMA" if head == "bf16" else "FP32 vector"})
    return out


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "knn_build_1Mx128_hbm_traffic.txt")


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 FETCH_SIZE/WRITE_SIZE summary (scaled
    by the copy-kernel calibration recorded there), or None."""
    try:
        for line in open(TRAFFIC_FILE):
            if line.startswith("traffic "):
                kv = dict(x.split("=", 1) for x in line.split()[1:])
                if kv["kernel"] == kernel:
                    return float(kv["traffic_bytes"])
    except OSError:
        pass
    return None


def roofline_of(stats, head, dim):
    """HBM roofline of the timed kernel family with the largest share of the build."""
    fams = [(sum(s["timed"][f]["ms"] for s in stats), f) for f in ("join", "margins", "leaf_dist", "leaf_push")]
    _, fam = max(fams)
    k = kernel_line(stats, fam, head, dim)
    if not k.get("launches"):
        return None
    return {"bound": "hbm", "achieved": k["achieved_gbs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": k["frac_hbm"], "traffic": measured_traffic(k["kernel"]), "kernel": k["kernel"],
            "avg_launch_ms": k["avg_launch_ms"], "bytes_per_launch": k["bytes_per_launch"]}


def precision_form(torch, ctx, L, args):
    """1,000,000 x 960 Gaussian mixture (100 centres U[0,1]^960, sigma 0.1, clipped to [0,1]), squared Euclidean,
    k=30 (max_candidates 30, leaf 150, 12 trees, <= 20 iterations), counter draws: the BF16-MFMA join with FP32
    re-evaluation inside the proven margin against the FP32 join.  Reports points/s of both, the fraction of
    pairs re-evaluated, recall@30 of both against an FP64 brute force (tie-aware) and of the BF16 graph against
    the FP32 graph, and the largest |stored - FP64| distance of the sampled rows."""
    from rinnd_amd import datasets
    n, d, k = args.pf_n, 960, 30
    log(f"[bench] precision form: generating {n}x{d} Gaussian mixture")
    X = datasets.cfg3(n=n, d=d)
    Xd = torch.from_numpy(X).cuda()
    del X
    pa = argparse.Namespace(**vars(args))
    pa.k, pa.trees, pa.leaf, pa.iters, pa.delta, pa.rng = k, 12, 150, 20, 0.001, "counter"
    outs = (torch.empty((n, k), dtype=torch.int32, device="cuda"),
            torch.empty((n, k), dtype=torch.float32, device="cuda"),
            torch.empty((n, k), dtype=torch.uint8, device="cuda"))
    res = {}
    for prec in ("fp32", "bf16"):
        gpu_build(ctx, L, Xd, n, d, pa, prec, outs)  # warm-up
        torch.cuda.synchronize()
        times, stats = [], []
        for s_ in range(args.pf_steps):
            t0 = time.perf_counter()
            st = gpu_build(ctx, L, Xd, n, d, pa, prec, outs)
            torch.cuda.synchronize()
            times.append(time.perf_counter() - t0)
            stats.append(st)
            log(f"[bench] precision form {prec} step {s_ + 1}: {times[-1] * 1e3:.1f} ms, iters {st['iterations']}, "
                f"forest {st['ms_forest']:.1f} init {st['ms_init']:.1f} cand {st['ms_candidates']:.1f} "
                f"join {st['ms_join']:.1f} apply {st['ms_apply']:.1f} ms, rescored {st['join_rescored']}")
        res[prec] = {"times": times, "stats": stats, "idx": outs[0].clone(), "dist": outs[1].clone()}
    rows = np.random.default_rng(3).choice(n, size=min(args.recall_rows, n), replace=False)
    rt = torch.as_tensor(rows, device="cuda")
    true_d, _ = exact_knn_f64(torch, Xd, rows, k, chunk=500)
    out = {"workload": f"{n}x{d} Gaussian mixture (100 centres, sigma 0.1, clipped to [0,1]), sqeuclidean, k={k}, "
                       f"max_candidates {k}, leaf 150, 12 trees, <= 20 iterations, counter draws",
           "dtype": "bf16-join/f32"}
    X64 = Xd.double()
    for prec, r in res.items():
        st = r["stats"][-1]
        it = st["iterations"]
        pairs = sum(st["join_pairs"][:it])
        resc = sum(st["join_rescored"][:it])
        ids = r["idx"][rt]
        f = ids.long().clamp(min=0)
        d64 = ((X64[f] - X64[rt][:, None, :]) ** 2).sum(-1)
        dd = (r["dist"][rt].double() - d64).abs()
        dd[ids < 0] = 0
        out[prec] = {"points_per_s": n / (sum(r["times"]) / len(r["times"])),
                     "ms_per_build": 1e3 * sum(r["times"]) / len(r["times"]), "iterations": it,
                     "phase_ms": {x: st[x] for x in ("ms_forest", "ms_init", "ms_candidates", "ms_join", "ms_apply")},
                     "rescored_fraction": (resc / pairs) if (prec == "bf16" and pairs) else None,
                     "recall_at_k_vs_fp64": tie_aware_recall(torch, Xd, Xd[rt], ids, true_d, k),
                     "max_abs_dist_err_vs_fp64": float(dd.max()),
                     "max_rel_dist_err_vs_fp64": float((dd / d64.clamp(min=1e-30)).max())}
    a, b = res["fp32"]["idx"][rt], res["bf16"]["idx"][rt]
    inter = (a[:, :, None] == b[:, None, :]).any(-1).double().sum(1) / k
    out["bf16"]["recall_at_k_vs_fp32_path"] = float(inter.mean())
    out["bf16_graph_equals_fp32_graph"] = bool(torch.equal(res["fp32"]["idx"], res["bf16"]["idx"]) and
                                               torch.equal(res["fp32"]["dist"].view(torch.int32),
                                                           res["bf16"]["dist"].view(torch.int32)))
    out["bf16_speedup"] = out["bf16"]["points_per_s"] / out["fp32"]["points_per_s"]
    log(f"[bench] precision form: fp32 {out['fp32']['ms_per_build']:.1f} ms, bf16 {out['bf16']['ms_per_build']:.1f} ms "
        f"({out['bf16_speedup']:.2f}x), rescored {out['bf16']['rescored_fraction']}, recall fp32 "
        f"{out['fp32']['recall_at_k_vs_fp64']:.4f} bf16 {out['bf16']['recall_at_k_vs_fp64']:.4f}, identical "
        f"{out['bf16_graph_equals_fp32_graph']}")
    del X64, Xd
    torch.cuda.empty_cache()
    return out


def batching_form(torch, ctx, L, args):
    """512 independent indexes, n log-uniform in [2,000, 100,000] (numpy seed 5), d=64 standard-Gaussian rows
    L2-normalised, Cosine, k=10 (max_candidates 10, leaf 60, delta 0.001; per index n_trees = clamp(round(2 log10 n),
    3, 12) and n_iters = max(ceil(log2 n), 5)), reference draws with one candidate thread.  One rinnd_nn_descent_batch
    call (work-queue forest, hipGraph-replayed passes, tuned join tiles) against the same indexes built by 512
    sequential rinnd_nn_descent calls; both from device-resident rows to row-sorted device heaps (sort_all
    included).  The two sets of heaps must be identical; recall@10 is sampled against an FP64 brute force."""
    import math
    from rinnd_amd import datasets
    m, d, k = args.bf_count, 64, 10
    sizes = datasets.cfg4_sizes(m)
    off = np.zeros(m + 1, np.int64)
    off[1:] = np.cumsum(sizes)
    N = int(off[-1])
    log(f"[bench] batching form: generating {m} indexes, {N} rows x {d}")
    X = np.empty((N, d), np.float32)
    for i, n in enumerate(sizes):
        X[off[i]:off[i + 1]] = datasets.cfg4_index(n, d, seed=5, i=i)
    Xd = torch.from_numpy(X).cuda()
    del X
    nt = np.array([min(12, max(3, int(round(2 * math.log10(n))))) for n in sizes], np.int64)
    ni = np.array([max(int(math.ceil(math.log2(n))), 5) for n in sizes], np.int64)
    seeds = [1000 + i for i in range(m)]
    p = L.make_params(k, 0, 60, k, 0, 0.001, angular=True, n_threads=1, rng_mode="reference")
    outs_b = (torch.empty((N, k), dtype=torch.int32, device="cuda"), torch.empty((N, k), dtype=torch.float32, device="cuda"),
              torch.empty((N, k), dtype=torch.uint8, device="cuda"))
    outs_s = tuple(torch.empty_like(t) for t in outs_b)
    its = np.zeros(m, np.int32)
    conv = np.zeros(m, np.int32)

    def batched():
        states = np.ascontiguousarray(np.stack([L.rng_seed(sd) for sd in seeds]), np.uint64)
        st = L.BatchStats()
        t0 = time.perf_counter()
        rc = L.lib().rinnd_nn_descent_batch(ctx.h, L.ptr(Xd), L.ptr(off), m, d, L.DIST["cosine"], C.byref(p),
                                            L.ptr(nt), L.ptr(ni), L.ptr(states), L.ptr(outs_b[0]), L.ptr(outs_b[1]),
                                            L.ptr(outs_b[2]), L.ptr(its), L.ptr(conv), 0, C.byref(st))
        ctx.check(rc, "rinnd_nn_descent_batch")
        ctx.check(L.lib().rinnd_sort_rows(ctx.h, L.ptr(outs_b[0]), L.ptr(outs_b[1]), L.ptr(outs_b[2]), N, k),
                  "rinnd_sort_rows")
        torch.cuda.synchronize()
        return time.perf_counter() - t0, st.to_dict()

    batched()  # warm-up (the join tile shape is tuned here and kept by the context)
    times, stats = [], []
    for s_ in range(args.bf_steps):
        t, st = batched()
        times.append(t)
        stats.append(st)
        log(f"[bench] batching form batched step {s_ + 1}: {t * 1e3:.1f} ms, passes {st['passes']}, graph replays "
            f"{st['graph_replays']}, forest {st['ms_forest']:.1f} init {st['ms_init']:.1f} cand+join "
            f"{st['ms_candidates_join']:.1f} apply {st['ms_apply']:.1f} ms, join shape "
            f"{st['join_threads_per_row']}x{st['join_rows_per_group']}")
    t_b = float(np.median(times))
    # the same indexes one rinnd_nn_descent call each
    row_b = d * 4

    def sequential():
        t0 = time.perf_counter()
        for i in range(m):
            a, n = int(off[i]), int(sizes[i])
            pi = L.make_params(k, int(nt[i]), 60, k, int(ni[i]), 0.001, angular=True, n_threads=1, rng_mode="reference")
            st = L.rng_seed(seeds[i])
            h = C.c_void_p()
            oi = C.c_void_p(outs_s[0].data_ptr() + a * k * 4)
            od = C.c_void_p(outs_s[1].data_ptr() + a * k * 4)
            of = C.c_void_p(outs_s[2].data_ptr() + a * k)
            rc = L.lib().rinnd_nn_descent(ctx.h, C.c_void_p(Xd.data_ptr() + a * row_b), n, d, L.DIST["cosine"],
                                          C.byref(pi), L.ptr(st), oi, od, of, C.byref(h), None)
            ctx.check(rc, "rinnd_nn_descent")
            L.lib().rinnd_forest_free(h)
            ctx.check(L.lib().rinnd_sort_rows(ctx.h, oi, od, of, n, k), "rinnd_sort_rows")
        torch.cuda.synchronize()
        return time.perf_counter() - t0

    if args.bf_no_sequential:
        t_s, same = float("nan"), None
    else:
        t_s = sequential()
        log(f"[bench] batching form sequential: {t_s * 1e3:.1f} ms for {m} builds")
        same = bool(torch.equal(outs_b[0], outs_s[0]) and
                    torch.equal(outs_b[1].view(torch.int32), outs_s[1].view(torch.int32)) and
                    torch.equal(outs_b[2], outs_s[2]))
    # recall@10 on 100 rows of each of 16 indexes, FP64 brute force within the index (1 - cos on unit rows)
    rng = np.random.default_rng(3)
    hits, total = 0.0, 0
    for i in rng.choice(m, size=min(16, m), replace=False):
        a, b = int(off[i]), int(off[i + 1])
        Xi = Xd[a:b].double()
        rows = torch.as_tensor(rng.choice(b - a, size=min(100, b - a), replace=False), device="cuda")
        D = 1.0 - Xi[rows] @ Xi.T
        D[torch.arange(len(rows), device="cuda"), rows] = float("inf")
        kth = torch.topk(D, k, dim=1, largest=False).values[:, k - 1:k]
        f = outs_b[0][a:b][rows].long()
        fd = torch.gather(D, 1, f.clamp(min=0))
        fd[f < 0] = float("inf")
        hits += float((fd <= kth + 1e-12).double().sum())
        total += len(rows) * k
    st = stats[-1]
    out = {"workload": f"{m} independent indexes, n log-uniform in [2000, 100000] (sum {N}), {d}-dim N(0,1) rows "
                       f"L2-normalised, Cosine, k={k}, max_candidates {k}, leaf 60, per-index n_trees clamp(round(2 log10 n), "
                       f"3, 12) and n_iters max(ceil(log2 n), 5), reference draws, one candidate thread",
           "dtype": "f32", "rows": N,
           "batched": {"points_per_s": N / t_b, "ms": 1e3 * t_b, "passes": st["passes"],
                       "graph_replays": st["graph_replays"], "graph_memsets_replaced": st["graph_memsets_replaced"],
                       "join_tile": {"threads_per_row": st["join_threads_per_row"],
                                     "rows_per_workgroup": st["join_rows_per_group"]},
                       "phase_ms": {x: st[x] for x in ("ms_forest", "ms_init", "ms_candidates_join", "ms_apply")}},
           "sequential": {"points_per_s": N / t_s, "ms": 1e3 * t_s, "builds": m},
           "speedup_vs_sequential": t_s / t_b,
           "heaps_identical_to_sequential": same,
           "iterations_range": [int(its.min()), int(its.max())], "converged": int(conv.sum()),
           "recall_at_10_sampled": hits / max(total, 1)}
    log(f"[bench] batching form: batched {t_b * 1e3:.1f} ms ({N / t_b / 1e6:.2f} M pts/s) vs sequential "
        f"{t_s * 1e3:.1f} ms ({N / t_s / 1e6:.2f} M pts/s): {t_s / t_b:.1f}x, identical {same}, recall "
        f"{out['recall_at_10_sampled']:.4f}")
    del Xd, outs_b, outs_s
    torch.cuda.empty_cache()
    return out


def serving_form(torch, ctx, L, args):
    """Resident 2,000,000 x 96 unit-norm index (500 centres N(0, I), sigma 0.2, rows normalised), cosine on
    normalised rows (DirectNormalizedCosine), graph k=20 (search degree <= 30); a stream of 4,000,000 queries from the
    same mixture, in arrival batches of 1-4,096 queries (log-uniform), k in {10, 100} (4:1), epsilon in {0.05, 0.2}
    (1:1), through the query service (rinnd_server_*): first saturated (4,096-query batches back to back), then an
    open-loop Poisson stream at a fraction of that rate, for the FP32 search, for FP8 e4m3 candidates on MFMA
    with FP32 re-ranking and for the int8 codes of the reference's quantized cosine index.  Latency of a batch =
    completion - scheduled arrival."""
    import threading
    from rinnd_amd import datasets
    sys.setswitchinterval(1e-4)  # the waiter thread takes the GIL promptly when a request completes
    n, d, nq = args.sv_n, 96, args.sv_queries
    log(f"[bench] serving form: generating {n}x{d} index rows and {nq} queries")
    X, centres = datasets.cfg5(n=n, d=d)
    Q = datasets.cfg5_queries(centres, nq)
    Xd = torch.from_numpy(X).cuda()
    Qd = torch.from_numpy(Q).cuda()
    del X
    p = L.IndexParams(20, 12, 1, 1, -1, -1, -1, 0.001, 1.0, 1.5, 1, 42, 0, args.threads, L.RNG["counter"],
                      L.JOIN["fp32"])
    h = C.c_void_p()
    stats = L.BuildStats()
    t0 = time.perf_counter()
    ctx.check(L.lib().rinnd_index_build(ctx.h, L.ptr(Xd), n, d, L.DIST["direct_cosine"], L.CORR["none"], C.byref(p),
                                        C.byref(h), C.byref(stats)), "rinnd_index_build")
    t_index = time.perf_counter() - t0
    log(f"[bench] serving form: index built in {t_index:.2f} s ({stats.iterations} iterations)")
    if "int8" in args.sv_modes.split(","):  # cosine_distance_mode="int8": per-vector codes of the rows
        t0 = time.perf_counter()
        ctx.check(L.lib().rinnd_index_encode_i8(h, 0), "rinnd_index_encode_i8")
        log(f"[bench] serving form: int8 codes in {time.perf_counter() - t0:.3f} s")
    # FP64 ground truth of a sample of the stream (cosine on unit rows: largest dot products)
    sample = np.arange(0, nq, max(1, nq // 10_000))
    X64 = Xd.double()
    true_dot = []
    for s0 in range(0, len(sample), 1000):
        qs = Qd[torch.as_tensor(sample[s0:s0 + 1000], device="cuda")].double()
        true_dot.append(torch.topk(qs @ X64.T, 10, dim=1).values)
    true_dot = torch.cat(true_dot)
    del X64
    rng = np.random.default_rng(9)
    sizes = []
    tot = 0
    while tot < nq:  # log-uniform arrival batches of 1..4096 queries
        b = int(min(nq - tot, max(1, round(float(np.exp(rng.uniform(0.0, np.log(4096.0))))))))
        sizes.append(b)
        tot += b
    ks = np.where(rng.random(len(sizes)) < 0.8, 10, 100)
    eps = np.where(rng.random(len(sizes)) < 0.5, 0.05, 0.2)
    starts = np.concatenate([[0], np.cumsum(sizes)[:-1]])
    Oi = torch.empty(nq * 100, dtype=torch.int32, device="cuda")
    Od = torch.empty(nq * 100, dtype=torch.float32, device="cuda")
    qbase, ibase, dbase = Qd.data_ptr(), Oi.data_ptr(), Od.data_ptr()
    out = {"workload": f"{n}x{d} unit-norm Gaussian mixture (500 centres N(0, I), sigma 0.2, normalised), "
                       f"DirectNormalizedCosine, graph k=20 "
                       f"(search degree <= 30), {nq} queries, arrival batches 1-4096 (log-uniform), k in {{10, 100}} "
                       f"(4:1), epsilon in {{0.05, 0.2}} (1:1)", "index_build_s": t_index, "batches": len(sizes),
           "server": {"max_batch": 4096, "max_inflight": args.sv_inflight, "target_batch": args.sv_target,
                      "visited_gib": args.sv_visited_gb,
                      "concurrent_searches": int(min(16384, args.sv_visited_gb * (1 << 30) // (n + 8 * 65536 + 4)))}}

    def widths(k, fp8):
        return (64 if k == 10 else 128, 64 if k == 10 else 128) if fp8 else (None, None)

    def run(srv, fp8, arrivals):
        """Submits batch b at arrivals[b] (seconds after the start, None: at once); returns the wall time, each
        batch's latency and the per-batch output offsets.  A second thread waits the tickets in order."""
        tickets = [None] * len(sizes)
        done_t = [0.0] * len(sizes)
        ready = threading.Semaphore(0)

        def waiter():
            for b in range(len(sizes)):
                ready.acquire()
                srv.wait(tickets[b])
                done_t[b] = time.perf_counter()
        th = threading.Thread(target=waiter)
        th.start()
        t_start = time.perf_counter()
        sched = [0.0] * len(sizes)
        for b, (st, sz) in enumerate(zip(starts, sizes)):
            if arrivals is not None:
                sched[b] = t_start + arrivals[b]
                gap = sched[b] - time.perf_counter()
                if gap > 3e-4:  # sleep (releasing the GIL for the waiter thread), then spin the last 0.3 ms
                    time.sleep(gap - 3e-4)
                while time.perf_counter() < sched[b]:
                    pass
            else:
                sched[b] = time.perf_counter()
            k = int(ks[b])
            cw, rw = widths(k, fp8)
            tickets[b] = srv.submit(qbase + int(st) * d * 4, sz, k, float(eps[b]), ibase + int(st) * 100 * 4,
                                    dbase + int(st) * 100 * 4, candidate_width=cw, rerank_width=rw)
            ready.release()
        th.join()
        wall = max(done_t) - t_start
        lat = np.array([done_t[b] - sched[b] for b in range(len(sizes))]) * 1e3
        return wall, lat

    def recall():
        """recall@10 of the sampled queries by (k, epsilon) class: a hit if its FP64 dot is >= the 10th true dot."""
        bidx = np.searchsorted(starts, sample, side="right") - 1
        off = torch.as_tensor((starts[bidx] * 100 + (sample - starts[bidx]) * ks[bidx]).astype(np.int64), device="cuda")
        ids = torch.stack([Oi[off + j] for j in range(10)], 1).long()
        qs = Qd[torch.as_tensor(sample, device="cuda")].double()
        dots = (Xd[ids.clamp(min=0)].double() * qs[:, None, :]).sum(-1)
        hit = ((dots >= true_dot[:, 9:10] - 1e-9) & (ids >= 0)).double().mean(1).cpu().numpy()
        res = {}
        for k in (10, 100):
            for e in (0.05, 0.2):
                m = (ks[bidx] == k) & (eps[bidx] == e)
                if m.any():
                    res[f"k{k}_eps{e}"] = float(hit[m].mean())
        res["all"] = float(hit.mean())
        return res

    for mode in args.sv_modes.split(","):
        fp8 = mode == "fp8"
        floor = int(mode[len("fp32_w"):]) if mode.startswith("fp32_w") else 0
        srv = L.Server(h, max_batch=4096, max_inflight=args.sv_inflight, max_k=128,
                       visited_bytes=int(args.sv_visited_gb * (1 << 30)),
                       quantized="int8" if mode == "int8" else fp8, target_batch=args.sv_target, min_width=floor)
        wall, lat = run(srv, fp8, None)  # saturated
        qps_max = nq / wall
        rec = recall()
        log(f"[bench] serving {mode}: saturated {qps_max:.0f} QPS, recall@10 {rec}")
        rate_q = args.sv_load * qps_max  # queries per second offered
        gaps = rng.exponential(1.0, len(sizes)) * (np.array(sizes) / rate_q)
        arrivals = np.cumsum(gaps) - gaps[0]
        wall_p, lat_p = run(srv, fp8, arrivals)
        st = srv.stats()
        srv.close()
        w = np.array(sizes, np.float64)
        order = np.argsort(lat_p)
        cw_ = np.cumsum(w[order]) / w.sum()  # query-weighted latency percentiles
        pq = lambda q: float(lat_p[order][np.searchsorted(cw_, q)])  # noqa: E731
        out[mode] = {"saturated_qps": qps_max, "recall_at_10": rec,
                     "poisson": {"offered_qps": rate_q, "achieved_qps": nq / wall_p,
                                 "p50_ms": pq(0.5), "p99_ms": pq(0.99), "batch_p50_ms": float(np.percentile(lat_p, 50)),
                                 "batch_p99_ms": float(np.percentile(lat_p, 99))},
                     "launches": st["launches"], "mean_launch_queries": st["queries"] / max(1, st["launches"]),
                     "gpu_ms": st["gpu_ms"]}
        if fp8:
            out[mode]["widths"] = "candidate/rerank 64 for k=10, 128 for k=100"
        if mode == "int8":
            out[mode]["widths"] = "candidate width k, no re-ranking (the reference's default quantized query)"
        if floor:
            out[mode]["min_width"] = floor
        log(f"[bench] serving {mode}: Poisson {nq / wall_p:.0f} QPS (offered {rate_q:.0f}), p50 {pq(0.5):.2f} ms, "
            f"p99 {pq(0.99):.2f} ms, {st['launches']} launches")
    L.lib().rinnd_index_free(h)
    del Xd, Qd, Oi, Od
    torch.cuda.empty_cache()
    return out


def defaults_build(torch, ctx, L, Xd, n, dim, args, outs, rows, true_d):
    """The same workload at the reference's default build parameters for n = 1,000,000: n_trees 12
    (clamp(round(2 log10 n), 3, 12)) and leaf_size max(60, min(256, 5k)) = 75 (index.rs:617-625), FP32 and BF16
    joins, the headline's draws and iteration cap otherwise."""
    a = argparse.Namespace(**vars(args))
    a.trees = min(12, max(3, int(round(2 * math.log10(n)))))
    a.leaf = max(60, min(256, 5 * args.k))
    out = {"n_trees": a.trees, "leaf_size": a.leaf}
    rt = torch.as_tensor(rows, device="cuda")
    for prec in ("fp32", "bf16"):
        gpu_build(ctx, L, Xd, n, dim, a, prec, outs)  # warm-up
        torch.cuda.synchronize()
        times = []
        for _ in range(max(1, min(args.steps, 3))):
            t0 = time.perf_counter()
            st = gpu_build(ctx, L, Xd, n, dim, a, prec, outs)
            torch.cuda.synchronize()
            times.append(time.perf_counter() - t0)
        t = float(np.median(times))
        rec = tie_aware_recall(torch, Xd, Xd[rt], outs[0][rt], true_d, args.k) if true_d is not None else None
        out[prec] = {"points_per_s": n / t, "ms_per_build": 1e3 * t, "recall_at_k": rec, "iterations": st["iterations"],
                     "phase_ms": {x: st[x] for x in ("ms_forest", "ms_init", "ms_candidates", "ms_join", "ms_apply")}}
        log(f"[bench] reference defaults ({a.trees} trees, leaf {a.leaf}) {prec}: {1e3 * t:.1f} ms, recall {rec}")
    out["bf16_speedup"] = out["bf16"]["points_per_s"] / out["fp32"]["points_per_s"]
    return out


def reference_build(torch, ctx, L, Xd, n, dim, args, outs, rows, true_d):
    """One FP32 build with the reference's FastRng streams (bit-exact mode)."""
    class A:
        pass
    a = A()
    a.__dict__.update(vars(args))
    a.rng = "reference"
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    st = gpu_build(ctx, L, Xd, n, dim, a, "fp32", outs)
    torch.cuda.synchronize()
    t = time.perf_counter() - t0
    rt = torch.as_tensor(rows, device="cuda")
    rec = tie_aware_recall(torch, Xd, Xd[rt], outs[0][rt], true_d, args.k) if true_d is not None else None
    log(f"[bench] reference-stream fp32 build: {t * 1e3:.1f} ms, recall@{args.k} {rec}")
    return {"points_per_s": n / t, "ms_per_build": t * 1e3, "recall_at_k": rec, "iterations": st["iterations"],
            "phase_ms": {x: st[x] for x in ("ms_forest", "ms_init", "ms_candidates", "ms_join", "ms_apply")}}


def query_bench(torch, L, ctx, X, Xd, args):
    from rinnd_amd import datasets
    n, dim = X.shape
    p = L.IndexParams(args.k, args.trees, 1, 1, args.leaf, min(args.k, 60), args.iters, 0.001, 1.0, 1.5, 0, 42, 0,
                      args.threads, L.RNG[args.rng], L.JOIN["fp32"])
    h = C.c_void_p()
    stats = L.BuildStats()
    t0 = time.perf_counter()
    ctx.check(L.lib().rinnd_index_build(ctx.h, L.ptr(Xd), n, dim, L.DIST["sqeuclidean"], L.CORR["sqrt"],
                                        C.byref(p), C.byref(h), C.byref(stats)), "rinnd_index_build")
    t_index = time.perf_counter() - t0
    log(f"[bench] index built in {t_index:.2f} s")
    rng = np.random.default_rng(8)
    Qn = np.clip(np.floor(X[rng.integers(0, n, args.queries)] + rng.normal(0, 2.0, (args.queries, dim)) + 0.5),
                 0, 255).astype(np.float32)
    Qd = torch.from_numpy(Qn).cuda()
    kq = 10
    oi = torch.empty((args.queries, kq), dtype=torch.int32, device="cuda")
    od = torch.empty((args.queries, kq), dtype=torch.float32, device="cuda")
    true_d, _ = exact_knn_queries_f64(torch, Xd, Qd, kq)
    one_i = np.empty(kq, np.int32)
    one_d = np.empty(kq, np.float32)
    runs = []
    # FP32 search at each epsilon, then the FP8 e4m3 candidate search (width 32) with exact re-ranking
    modes = [(float(e), None) for e in args.epsilons.split(",")]
    modes += [(float(e), 32) for e in args.fp8_epsilons.split(",") if e]

    def run(nq_h, q_h, k_, eps, width, out_i, out_d):
        if width is None:
            return L.lib().rinnd_index_query(h, q_h, nq_h, k_, C.c_float(eps), out_i, out_d)
        return L.lib().rinnd_index_query_quantized(h, q_h, nq_h, k_, C.c_float(eps), width, width, out_i, out_d)

    for eps, width in modes:
        name = "fp32" if width is None else f"fp8+rerank{width}"
        for _ in range(2):
            ctx.check(run(args.queries, L.ptr(Qd), kq, eps, width, L.ptr(oi), L.ptr(od)), name)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        reps = 3
        for _ in range(reps):
            ctx.check(run(args.queries, L.ptr(Qd), kq, eps, width, L.ptr(oi), L.ptr(od)), name)
        torch.cuda.synchronize()
        qps = args.queries * reps / (time.perf_counter() - t0)
        rec = tie_aware_recall(torch, Xd, Qd, oi, true_d, kq)
        lat = []
        for j in range(200):
            q = np.ascontiguousarray(Qn[j])
            t0 = time.perf_counter()
            ctx.check(run(1, L.ptr(q), kq, eps, width, L.ptr(one_i), L.ptr(one_d)), name)
            lat.append(time.perf_counter() - t0)
        lat_ms = np.array(lat) * 1e3
        log(f"[bench] query {name} eps {eps}: {qps:.0f} QPS, recall@10 {rec:.4f}, "
            f"p50 {np.percentile(lat_ms, 50):.3f} ms, p99 {np.percentile(lat_ms, 99):.3f} ms")
        runs.append({"search": name, "qps": qps, "batch": args.queries, "k": kq, "epsilon": eps,
                     "recall_at_10": rec, "p50_ms_single": float(np.percentile(lat_ms, 50)),
                     "p99_ms_single": float(np.percentile(lat_ms, 99))})
    L.lib().rinnd_index_free(h)
    ok = [r for r in runs if r["recall_at_10"] >= 0.95]
    best = max(ok, key=lambda r: r["qps"]) if ok else max(runs, key=lambda r: r["recall_at_10"])
    # the chosen setting on indexes built from two more seeds (QPS varies with the graph's random draws)
    seeds = [{"seed": 42, "qps": best["qps"], "recall_at_10": best["recall_at_10"]}]
    width = None if best["search"] == "fp32" else int(best["search"].split("rerank")[1])
    for seed in (43, 44):
        p.random_seed = seed
        h = C.c_void_p()
        ctx.check(L.lib().rinnd_index_build(ctx.h, L.ptr(Xd), n, dim, L.DIST["sqeuclidean"], L.CORR["sqrt"],
                                            C.byref(p), C.byref(h), C.byref(stats)), "rinnd_index_build")
        for _ in range(2):
            ctx.check(run(args.queries, L.ptr(Qd), kq, best["epsilon"], width, L.ptr(oi), L.ptr(od)), "query")
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(3):
            ctx.check(run(args.queries, L.ptr(Qd), kq, best["epsilon"], width, L.ptr(oi), L.ptr(od)), "query")
        torch.cuda.synchronize()
        seeds.append({"seed": seed, "qps": args.queries * 3 / (time.perf_counter() - t0),
                      "recall_at_10": tie_aware_recall(torch, Xd, Qd, oi, true_d, kq)})
        L.lib().rinnd_index_free(h)
        log(f"[bench] query {best['search']} eps {best['epsilon']} index seed {seed}: {seeds[-1]['qps']:.0f} QPS, "
            f"recall@10 {seeds[-1]['recall_at_10']:.4f}")
    q = [x["qps"] for x in seeds]
    return dict(best, index_build_s=t_index, sweep=runs, index_seeds=seeds, qps_mean_over_seeds=float(np.mean(q)),
                qps_min_over_seeds=float(np.min(q)))


if __name__ == "__main__":
    main()
