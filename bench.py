"""Benchmark of the exact kNN join on one MI355X.

    python bench.py [--config c2] [--steps K] [--warmup W] [--gpus 1]

A step is one whole exact join (index build + bounds + join + rerank) of
the configuration's query set against its reference set, inputs already
resident in HBM.  `value` is queries per second over the K timed steps.
Workloads (seeded synthetic data, knn_sweet_amd.datagen):
  c1  self-join of 245,057 integer points, d=3, k=200
  c2  |R|=1,000,000, |Q|=100,000, d=64, k=32, FP32 Gaussian mixture
  c3  |R|=2,000,000, |Q|=200,000, d=128, k=100, FP32 Gaussian mixture
The roofline object describes the join kernel (the dominant one); its
duration is the hipEvent time the library records around it on its own
stream.  cpu_baseline times the FP64 brute-force oracle on a bounded sample
of the same queries on this host's cores.
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

METRIC = "exact kNN queries/s/GPU (BF16-filter vs plain FP32, recall=1.0); p99 ms"
HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3
BF16_PEAK_TFLOPS = 2500.0


def load_config(name: str):
    from knn_sweet_amd import datagen
    if name == "c1":
        X = datagen.skin_like(245_057, seed=2016)
        return dict(workload="selfjoin-int-d3-245057-k200", R=X, Q=X, k=200, self_join=True)
    if name == "c2":
        R, Q = datagen.mixture64(1_000_000, 100_000, seed=2017)
        return dict(workload="mixture-d64-R1M-Q100k-k32", R=R, Q=Q, k=32, self_join=False)
    if name == "c3":
        R, Q = datagen.mixture128(2_000_000, 200_000, seed=2018)
        return dict(workload="mixture-d128-R2M-Q200k-k100", R=R, Q=Q, k=100, self_join=False)
    raise SystemExit(f"unknown config {name}")


def cpu_baseline(R, Q, k, target_s: float = 12.0):
    """FP64 brute-force oracle on a bounded query sample (calibrated to ~target_s)."""
    import oracle
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or (os.cpu_count() or 1)
    threads = min(threads, 16)
    n = min(32, Q.shape[0])
    t0 = time.perf_counter()
    oracle.bruteforce(R, Q[:n], k, nthreads=threads)
    dt = time.perf_counter() - t0
    n2 = int(min(Q.shape[0], max(n, n * target_s / max(dt, 1e-3))))
    if n2 > n:
        t0 = time.perf_counter()
        oracle.bruteforce(R, Q[:n2], k, nthreads=threads)
        dt = time.perf_counter() - t0
        n = n2
    return {"value": n / dt, "unit": "queries/s", "cores": threads, "kind": "port",
            "sample": f"first {n} queries of the same workload, FP64 brute force (oracle), {dt:.1f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", default="c2")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--brute-force", action="store_true", help="one cluster per side")
    ap.add_argument("--kernel", default="auto", choices=["auto", "valu", "mfma"])
    ap.add_argument("--precision", default="auto", choices=["auto", "fp32", "bf16", "bf16x1"])
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1"}))
        sys.exit(2)

    import torch
    import knn_sweet_amd as kj

    cfg = load_config(args.config)
    R, Q, k = cfg["R"], cfg["Q"], cfg["k"]
    Rd = torch.from_numpy(R).cuda()
    Qd = Rd if cfg["self_join"] else torch.from_numpy(Q).cuda()
    torch.cuda.synchronize()
    ctx = kj.Context(0)
    opts = dict(brute_force=args.brute_force,
                kernel={"auto": kj.KERNEL_AUTO, "valu": kj.KERNEL_VALU, "mfma": kj.KERNEL_MFMA}[args.kernel],
                precision={"auto": kj.PREC_AUTO, "fp32": kj.PREC_FP32, "bf16": kj.PREC_BF16,
                           "bf16x1": kj.PREC_BF16X1}[args.precision])
    for _ in range(args.warmup):
        ctx.join(Rd, Qd, k, **opts)
    torch.cuda.synchronize()
    times, stats = [], []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        s0 = time.perf_counter()
        _, _, st = ctx.join(Rd, Qd, k, **opts)
        times.append((time.perf_counter() - s0) * 1e3)
        stats.append(st)
    torch.cuda.synchronize()
    total = time.perf_counter() - t0
    nq, nr, d = Q.shape[0], R.shape[0], R.shape[1]
    value = nq * args.steps / total

    t_join = float(np.mean([s["t_join_ms"] for s in stats]))
    evals = float(np.mean([s["n_dist_evals"] for s in stats]))
    st = stats[-1]
    if st["exact_mode"]:
        b_alg = 4.0 * d * (nq + nr) + 8.0 * k * nq
        achieved = b_alg / (t_join * 1e-3) / 1e9
        roof = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS, "traffic": None}
    else:
        # executed flops per (query, candidate) pair: 3d for the direct form
        # (d subtractions, d FMAs), 2d for the GEMM form on the matrix cores
        mfma = st["kernel_used"] == kj.KERNEL_MFMA
        flops = evals * (2.0 if mfma else 3.0) * d
        peak = BF16_PEAK_TFLOPS if st["precision_used"] in (kj.PREC_BF16, kj.PREC_BF16X1) else FP32_PEAK_TFLOPS
        achieved = flops / (t_join * 1e-3) / 1e12
        roof = {"bound": "mfma", "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
                "frac": achieved / peak, "traffic": None}
    out = {
        "metric": METRIC, "value": value, "unit": "queries/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": total * 1e3 / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded Gaussian mixture, knn_sweet_amd.datagen)",
        "config": {"workload": cfg["workload"], "n_r": nr, "n_q": nq, "d": d, "k": k,
                   "precision": {1: "fp32", 2: "bf16x3-filter+fp64-rerank",
                                 4: "bf16x1-filter+fp64-rerank"}[st["precision_used"]],
                   "kernel": {1: "valu", 2: "mfma"}[st["kernel_used"]],
                   "self_join": cfg["self_join"]},
        "p50_ms": float(np.percentile(times, 50)), "p99_ms": float(np.percentile(times, 99)),
        "roofline": roof,
        "stages_ms": {"index": st["t_index_ms"], "bounds": st["t_bounds_ms"], "join": t_join,
                      "fallback": st["t_fallback_ms"], "total": st["t_total_ms"]},
        "pruning_ratio": st["pruning_ratio"], "dist_evals": int(evals),
        "dist_pairs_per_s": evals / (t_join * 1e-3), "m_q": st["m_q"], "m_r": st["m_r"],
        "kcap": st["kcap"], "tile_q": st["tile_q"], "n_overflow": st["n_overflow"],
        "exact_mode": st["exact_mode"], "merges": st["n_merges"], "appended": st["n_appended"],
        "filter_max_abs_err": st["filter_max_abs_err"],
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(R, Q, k)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
