"""Harmony throughput benchmark on one MI355X.

Default workload: 1,300,000 cells x 50 PCs from a 50-type Gaussian mixture
with 16 uniform batches (one-hot Phi, B = 16) and per-batch shifts
~N(0, 0.5^2); K = 100, theta = 2, sigma = 0.1, harmony1 with lambda = 1,
block_proportion = 0.05, 20 clustering iterations per Harmony iteration (early
stop off).  One step = one Harmony iteration (clustering + ridge correction)
on device-resident state; the metric is cells x Harmony iterations per second.

One run measures, on the same workload:
  * plain FP32 and split-BF16 throughput (timed steps), and the batched form
    (several instances of the workload in one multi-job clustering loop);
    ``value`` is the split-BF16 batched figure, the others sit beside it;
  * the relative error of Z_corr, R and Y of both precisions after two Harmony
    iterations against the FP64 path from the same initial centroids;
  * p50/p99 job latency of a short served leg (Poisson job stream);
  * the roofline of the dominant kernel (algorithmic bytes / live HIP-event
    launch time vs 8 TB/s);
  * a CPU baseline: the NumPy FP64 oracle on this host, at the full N for two
    clustering iterations (median of 3) plus the correction on a sample,
    extrapolated to one Harmony iteration.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "Harmony cells·iters/s (split-BF16 vs FP32, ≤1e-3 rel err) + p50/p99 job latency"
HBM_PEAK_GBS = 8000.0
COPY_BYTES = 256 << 20

WORKLOADS = {
    # name: (n_cells, n_pcs, n_types, n_batches, n_clusters, Dirichlet batch sizes, shift sd)
    "c2": (200_000, 50, 30, 8, 100, True, 1.0),
    "c3": (1_300_000, 50, 50, 16, 100, False, 0.5),
}


def make_data(name: str, seed: int = 0):
    from rapids_singlecell_amd.datasets import synthetic_pcs

    n, d, types, B, K, dirichlet, shift = WORKLOADS[name]
    X, obs = synthetic_pcs(n, d, types, B, dirichlet_batches=dirichlet, seed=seed,
                           shift_scale=shift)
    codes = np.asarray(obs["batch"].cat.codes.values, dtype=np.int32)
    return X, codes, B, K


def workload_text(name: str) -> str:
    n, d, types, B, K, dirichlet, shift = WORKLOADS[name]
    sizes = "Dirichlet(1)-sized" if dirichlet else "uniform"
    return (f"harmony1 N={n} d={d} K={K} B={B} (one-hot, {sizes} batches, shifts sd {shift}) "
            f"{types}-type mixture, theta=2 sigma=0.1 lambda=1 block_proportion=0.05, "
            f"20 clustering iters per Harmony iter")


def initial_centroids(X, K, seed=7):
    """K random rows of the L2-normalised input (FP64): the same initial
    centroids for every precision of one comparison."""
    Xd = X.astype(np.float64)
    Z = Xd / np.maximum(np.linalg.norm(Xd, axis=1, keepdims=True), 1e-12)
    return Z[np.random.default_rng(seed).choice(len(X), K, replace=False)]


def build_job(X, codes, B, K, precision, stream=None):
    import torch

    from rapids_singlecell_amd.preprocessing._harmony._job import HarmonyJob

    dev = torch.device("cuda", 0)
    dt = X.dtype
    N_b = np.bincount(codes, minlength=B).astype(dt)
    job = HarmonyJob(torch.as_tensor(X, device=dev), codes, B, n_clusters=K,
                     theta_np=np.full(B, 2.0, dt), N_b_np=N_b, sigma=0.1,
                     block_proportion=0.05, max_iter_clustering=20,
                     tol_clustering=float("-inf"), ridge_lambda=1.0, stabilized_penalty=False,
                     dynamic_lambda=False, correction_method="batched", random_state=0,
                     precision=precision, stream=stream)
    return job


def time_steps(job, first_iter: int, steps: int) -> float:
    import torch

    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(steps):
        job.step(first_iter + i)
    torch.cuda.synchronize()
    return time.perf_counter() - t0


def roofline(job, first_iter: int, steps: int, esz: int, precision: str) -> tuple[dict, dict]:
    """Live HIP-event timing of each kernel kind over `steps` Harmony iterations."""
    from rapids_singlecell_amd import _native

    _native.profile_collect()
    _native.profile_enable(True)
    try:
        for i in range(steps):
            job.step(first_iter + i)
        prof = _native.profile_collect()
    finally:
        _native.profile_enable(False)
    n, d, K = job.n_cells, job.n_pcs, job.n_clusters
    nblk = -(-n // job.block_size)
    n_clust = steps * job.max_iter_clustering
    # algorithmic bytes per launch (reads and writes each array once)
    bytes_per = {
        "row": esz * (n / nblk) * (K + d),          # read Z rows, write R rows
        "sweep": esz * n * (K + d),                 # all blocks: read Z rows, write R rows
        "centroid": esz * n * (K + d),              # read R and Z
        "phi": esz * n * (K + d),                   # read R and X
        "apply": esz * n * (K + 3 * d),             # read R, X; write Z_hat, Z_norm
    }
    kernels = {}
    for kind, (ms, cnt) in prof.items():
        if cnt == 0:
            continue
        entry = {"launches": int(cnt), "total_ms": round(ms, 4), "avg_us": round(1e3 * ms / cnt, 3)}
        if kind in bytes_per:
            gbs = bytes_per[kind] / (ms / cnt * 1e-3) / 1e9
            entry["algorithmic_bytes_per_launch"] = int(bytes_per[kind])
            entry["achieved_GBs"] = round(gbs, 1)
        kernels[kind] = entry
    timed = {k: v for k, v in kernels.items() if "achieved_GBs" in v}
    dom = max(timed, key=lambda k: timed[k]["total_ms"])
    ach = timed[dom]["achieved_GBs"]
    traffic = None
    tfile = ROOT / "profiles" / "traffic.json"
    if tfile.exists():
        try:
            # HBM bytes per launch from rocprofv3 FETCH_SIZE + WRITE_SIZE passes,
            # each scaled by what the counter reads for a copy of known size
            traffic = json.loads(tfile.read_text()).get(f"{dom}:{n}x{d}x{K}:{esz}:{precision}")
        except Exception:
            traffic = None
    roof = {"bound": "hbm", "kernel": f"{dom}_kernel", "achieved": ach, "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": traffic}
    kernels["_per_clustering_iter_ms"] = round(
        sum(v["total_ms"] for k, v in kernels.items()
            if k in ("row", "sweep", "centroid", "yreduce")) / max(n_clust, 1), 4)
    # the block order is built on a side stream, overlapped with the iteration before
    return roof, kernels


def cpu_baseline(X, codes, B, K, reps: int = 3, n_clust: int = 2,
                 corr_sample: int = 25_000) -> dict:
    """NumPy FP64 oracle on this host, extrapolated to one Harmony iteration.

    Clustering: ``n_clust`` clustering iterations over all N cells (median of
    ``reps`` runs), scaled to 20.  Correction: the oracle's batched correction
    (whose apply step builds a K x n x d temporary) on the first
    ``corr_sample`` cells (median of ``reps``), scaled linearly to N.
    """
    import statistics

    from oracle import harmony_oracle as ho

    try:
        from threadpoolctl import threadpool_info

        blas_threads = max((i.get("num_threads", 1) for i in threadpool_info()), default=1)
    except Exception:  # noqa: BLE001
        blas_threads = None
    n = X.shape[0]
    Xd = X.astype(np.float64)
    cats = np.asarray(codes, dtype=np.int32)
    N_b = np.bincount(cats, minlength=B).astype(np.float64)
    Pr_b = (N_b / n).reshape(-1, 1)
    theta = np.full(B, 2.0)
    Z = ho.l2_normalize(Xd)
    Y0 = initial_centroids(X, K)
    R0, E0, O0, _ = ho.init_clusters(Z, Y0, sigma=0.1, Pr_b=Pr_b, theta=theta, cats=cats,
                                     n_batches=B, stabilized=False)
    tc = []
    for _ in range(reps):
        R, E, O = R0.copy(), E0.copy(), O0.copy()
        t0 = time.perf_counter()
        ho.clustering_loop(Z, R, E, O, Pr_b, cats, theta, sigma=0.1, tol=float("-inf"),
                           max_iter=n_clust, block_size=int(n * 0.05), seed=0, stabilized=False)
        tc.append(time.perf_counter() - t0)
    m = min(corr_sample, n)
    lam = ho.compute_lambda_kb(E, O, N_b, alpha=0.2, threshold=None, ridge_lambda=1.0,
                               dynamic_lambda=False)
    tk = []
    for _ in range(reps):
        t0 = time.perf_counter()
        ho.correction_batched(Xd[:m], R[:m], O, lam, cats[:m], B)
        tk.append(time.perf_counter() - t0)
    t_clust = statistics.median(tc)
    t_corr = statistics.median(tk)
    t_iter = t_clust * (20 / n_clust) + t_corr * (n / m)
    try:
        aff = len(os.sched_getaffinity(0))
    except Exception:  # noqa: BLE001
        aff = os.cpu_count()
    return {"value": round(n / t_iter, 1), "unit": "cells·iters/s",
            "cores": blas_threads or aff, "kind": "port",
            "sample": f"oracle FP64 (NumPy): {n_clust} clustering iterations over all {n} cells, "
                      f"median of {reps} = {t_clust:.2f} s, scaled x{20 / n_clust:g}; batched "
                      f"correction on the first {m} cells, median of {reps} = {t_corr:.3f} s, "
                      f"scaled x{n / m:.1f}; one Harmony iteration = {t_iter:.1f} s "
                      f"(os.cpu_count={os.cpu_count()}, affinity={aff}, "
                      f"blas_threads={blas_threads})"}


def served(args) -> int:
    out = run_served(args, args.served_duration)
    print(json.dumps(out), flush=True)
    return 0


def run_served(args, duration: float) -> dict:
    """Served form: a Poisson stream of integration jobs through HarmonyServer.

    Job shapes: N log-uniform in [5k, 500k], d in {30, 50, 100}, B uniform in
    [2, 32], theta=2, sigma=0.1, K = min(100, N/30), Harmony defaults (early
    stopping on), half FP32 and half split-BF16.  Datasets are drawn from a
    pool generated before the clock starts.  Reports p50/p99 job latency and
    the aggregate cells x Harmony iterations per second over the run.
    """
    import torch

    from rapids_singlecell_amd import _native
    from rapids_singlecell_amd.datasets import synthetic_pcs
    from rapids_singlecell_amd.serve import HarmonyServer, latency_summary

    _native.load()
    rng = np.random.default_rng(args.served_seed)
    pool = []
    for i in range(args.served_pool):
        n = int(np.exp(rng.uniform(np.log(5_000), np.log(500_000))))
        d = int(rng.choice([30, 50, 100]))
        B = int(rng.integers(2, 33))
        X, obs = synthetic_pcs(n, d, int(rng.integers(5, 41)), B, seed=1000 + i)
        codes = np.asarray(obs["batch"].cat.codes.values, dtype=np.int32)
        pool.append((X, codes, len(obs["batch"].cat.categories)))
    print(f"[bench] served: pool of {len(pool)} datasets ready", file=sys.stderr, flush=True)
    with HarmonyServer(slots=args.served_slots) as srv:
        # warm-up: one small job per precision
        for prec in ("native", "split_bf16"):
            X, c, B = min(pool, key=lambda p: p[0].shape[0])
            srv.submit(X, c, B, precision=prec, max_iter_harmony=2).result()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        futs, t = [], 0.0
        while True:
            t += rng.exponential(1.0 / args.served_rate)
            if t > duration:
                break
            now = time.perf_counter() - t0
            if t > now:
                time.sleep(t - now)
            X, c, B = pool[int(rng.integers(len(pool)))]
            prec = "split_bf16" if rng.random() < 0.5 else "native"
            futs.append((srv.submit(X, c, B, precision=prec, theta=2.0, sigma=0.1),
                         (X.shape[0], X.shape[1], B, prec)))
        res, failures = [], []
        for f, shape in futs:
            try:
                res.append(f.result())
            except Exception as e:  # noqa: BLE001 - reported in the output line
                failures.append({"shape": shape, "error": str(e)[:200]})
        wall = time.perf_counter() - t0
        stats = dict(srv.stats)
    lats = [info["latency_s"] for _, info in res]
    cells_iters = sum(info["n_cells"] * info["harmony_iters"] for _, info in res)
    out = {
        "metric": METRIC + " (served form)",
        "value": round(cells_iters / wall, 1),
        "unit": "cells·iters/s",
        "n_gpus": 1,
        "higher_is_better": True,
        "latency": latency_summary(lats),
        "jobs": len(res),
        "offered_rate_jobs_per_s": args.served_rate,
        "duration_s": duration,
        "wall_s": round(wall, 2),
        "slots": args.served_slots,
        "server_stats": stats,
        "failures": failures[:10],
        "n_failures": len(failures),
        "data": "synthetic Gaussian-mixture PCA embeddings (pool of "
                f"{len(pool)} shapes drawn from the served-form distribution)",
    }
    return out


def batched(args) -> int:
    """Batched form: many independent Harmony instances advanced together.

    Per instance: N = floor(exp(U(ln 2000, ln 100000))), d in {30, 50, 100},
    B uniform in [2, 32], K = max(2, floor(min(100, N/30))), 5-40-type
    mixtures with batch shifts, 10 Harmony x 20 clustering iterations (early
    stop off), initial centroids = K random rows.  ``harmonize_many`` runs the
    instances of one kernel shape through one multi-job clustering call per
    Harmony iteration.  The value is the total cells x Harmony iterations over
    the wall time of the whole set (inputs already on the device, from the
    first initialisation to the last correction).
    """
    import torch

    from rapids_singlecell_amd import _native
    from rapids_singlecell_amd.datasets import synthetic_pcs
    from rapids_singlecell_amd.preprocessing._harmony._batch import harmonize_many

    _native.load()
    rng = np.random.default_rng(args.batched_seed)
    probs = []
    for i in range(args.batched_n):
        n = int(np.exp(rng.uniform(np.log(2_000), np.log(100_000))))
        d = int(rng.choice([30, 50, 100]))
        B = int(rng.integers(2, 33))
        K = max(2, int(min(100, n / 30)))
        X, obs = synthetic_pcs(n, d, int(rng.integers(5, 41)), B, seed=5000 + i)
        Y0 = X[rng.choice(n, K, replace=False)]
        probs.append(dict(X=torch.as_tensor(X, device="cuda"), obs=obs, key="batch",
                          n_clusters=K, init_centroids=Y0, max_iter_harmony=10,
                          max_iter_clustering=20, tol_harmony=float("-inf"),
                          tol_clustering=float("-inf"), precision=args.precision,
                          random_state=i))
    print(f"[bench] batched: {len(probs)} instances ready", file=sys.stderr, flush=True)
    harmonize_many(probs[:2])  # warm-up (library load, kernel first launches)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    out_z = harmonize_many(probs)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    assert all(torch.isfinite(z).all().item() for z in out_z)
    cells_iters = sum(p["X"].shape[0] * 10 for p in probs)
    out = {
        "metric": METRIC + " (batched form)",
        "value": round(cells_iters / wall, 1),
        "unit": "cells·iters/s",
        "n_gpus": 1,
        "higher_is_better": True,
        "instances": len(probs),
        "total_cells": int(sum(p["X"].shape[0] for p in probs)),
        "wall_s": round(wall, 3),
        "precision": args.precision,
        "data": "synthetic Gaussian-mixture PCA embeddings, one per instance",
    }
    print(json.dumps(out), flush=True)
    return 0


def batched_leg(X, codes, B, K, precision: str, copies: int, warmup: int, steps: int) -> dict:
    """``copies`` instances of the workload (same data, different random
    states, so different block orders) stepped together: one multi-job
    clustering call per Harmony iteration, then every correction."""
    import torch

    from rapids_singlecell_amd.preprocessing._harmony._batch import step_many
    from rapids_singlecell_amd.preprocessing._harmony._job import HarmonyJob

    prec = {"native": 0, "split_bf16": 1}[precision]
    stream = torch.cuda.current_stream()
    dev = torch.device("cuda", 0)
    Xd = torch.as_tensor(X, device=dev)
    N_b = np.bincount(codes, minlength=B).astype(np.float32)
    jobs = []
    for c in range(copies):
        job = HarmonyJob(Xd, codes, B, n_clusters=K, theta_np=np.full(B, 2.0, np.float32),
                         N_b_np=N_b, sigma=0.1, block_proportion=0.05, max_iter_clustering=20,
                         tol_clustering=float("-inf"), ridge_lambda=1.0,
                         stabilized_penalty=False, dynamic_lambda=False,
                         correction_method="batched", random_state=c, precision=prec,
                         stream=stream)
        job.init()
        jobs.append(job)
    torch.cuda.synchronize()
    it = 0
    for _ in range(warmup):
        step_many(jobs, it)
        it += 1
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        step_many(jobs, it)
        it += 1
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    n = X.shape[0]
    leg = {"precision": precision, "instances": copies,
           "value": round(copies * n * steps / dt, 1),
           "ms_per_step": round(1e3 * dt / steps, 3), "steps": steps}
    print(f"[bench] batched {precision} x{copies}: {steps} steps in {dt:.3f} s -> "
          f"{leg['value']:.4g} cells·iters/s", file=sys.stderr, flush=True)
    del jobs
    torch.cuda.empty_cache()
    return leg


def multi_bench(args) -> int:
    """Several covariates (the precision-form configuration): 1.3M cells x 50
    PCs, 50-type mixture, two independent uniform covariates (16 batches and
    64 donors, 80 levels of stacked one-hot Phi), per-level shifts
    ~N(0, 0.5^2), K = 100, harmony1 lambda = 1.  Times Harmony iterations of
    the multi-covariate path in FP32, split-BF16 and FP64 (inputs in HBM) and
    reports the error of the FP32 paths against the FP64 path after two
    Harmony iterations from the same initial centroids."""
    import torch

    from rapids_singlecell_amd import _native
    from rapids_singlecell_amd.preprocessing._harmony._helper import _get_covariates
    from rapids_singlecell_amd.preprocessing._harmony._job import HarmonyJob

    _native.load()
    tune = {k: int(v) for k, v in (t.split("=", 1) for t in args.tune)}
    if tune:
        _native.set_tuning(**tune)
    n, d, K = args.multi_cells, 50, 100
    rng = np.random.default_rng(0)
    centers = rng.normal(0, 3.0, size=(50, d))
    b = rng.integers(0, 16, n)
    dn = rng.integers(0, 64, n)
    X = (centers[rng.integers(0, 50, n)] + rng.normal(size=(n, d))
         + rng.normal(0, 0.5, size=(16, d))[b] + rng.normal(0, 0.5, size=(64, d))[dn])
    X = X.astype(np.float32)
    import pandas as pd

    obs = pd.DataFrame({"batch": pd.Categorical(b), "donor": pd.Categorical(dn)})
    _, n_levels, group_levels, groups = _get_covariates(obs, ["batch", "donor"])
    N_b = np.bincount(group_levels[groups].ravel(), minlength=sum(n_levels))
    Y0 = initial_centroids(X, K)

    def make(Xh, prec):
        dt = Xh.dtype
        job = HarmonyJob(torch.as_tensor(Xh, device="cuda"), groups, len(group_levels),
                         n_clusters=K, theta_np=np.full(sum(n_levels), 2.0, dt),
                         N_b_np=N_b.astype(dt), sigma=0.1, block_proportion=0.05,
                         max_iter_clustering=20, tol_clustering=float("-inf"), ridge_lambda=1.0,
                         stabilized_penalty=False, dynamic_lambda=False, random_state=0,
                         precision=prec, group_levels=group_levels, n_levels0=n_levels[0])
        job.init(torch.as_tensor(Y0.astype(dt)))
        return job

    legs = {}
    final = {}
    for name, Xh, prec in (("fp32", X, 0), ("split_bf16", X, 1), ("fp64", X.astype(np.float64), 0)):
        job = make(Xh, prec)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        job.step(0)
        job.step(1)
        torch.cuda.synchronize()
        dt_ = (time.perf_counter() - t0) / 2
        legs[name] = {"value": round(n / dt_, 1), "ms_per_step": round(1e3 * dt_, 2)}
        final[name] = job.Z_hat.double()
        print(f"[bench] multi {name}: {1e3 * dt_:.1f} ms per Harmony iteration", file=sys.stderr,
              flush=True)
        del job
        torch.cuda.empty_cache()
    ref = final["fp64"]
    errs = {k: float(torch.linalg.norm(final[k] - ref) / torch.linalg.norm(ref))
            for k in ("fp32", "split_bf16")}
    print(json.dumps({
        "metric": "Harmony cells·iters/s, several covariates (stacked one-hot Phi)",
        "unit": "cells·iters/s", "higher_is_better": True,
        "config": {"workload": f"N={n} d={d} K={K} covariates 16 batches + 64 donors "
                               f"({len(group_levels)} groups, 80 levels), harmony1 lambda=1, "
                               "20 clustering iters per Harmony iter"},
        "legs": legs, "rel_err_Z_corr_vs_fp64_after_2_iters": errs,
    }), flush=True)
    return 0


def knn_bench(args) -> int:
    """Brute-force kNN (the neighbour search on the integrated embedding):
    every row of an N x 50 embedding against all rows, k = 15, euclidean.
    Timed with HIP events around ``hz_knn`` (inputs on the device); the
    roofline is the matrix-core bound, 2 N^2 d useful flops against the FP32
    MFMA peak."""
    import ctypes

    import torch

    from rapids_singlecell_amd import _native
    from rapids_singlecell_amd.datasets import synthetic_pcs

    lib = _native.load()
    out = []
    for n in args.knn_sizes:
        X, _ = synthetic_pcs(n, 50, 30, 8, seed=1)
        Xd = torch.as_tensor(X, device="cuda")
        k = 15
        idx = torch.empty((n, k), dtype=torch.int32, device="cuda")
        dist = torch.empty((n, k), dtype=torch.float32, device="cuda")
        nb = lib.hz_knn_workspace_bytes(n, 50, 0)
        ws = torch.empty(max(nb, 256), dtype=torch.uint8, device="cuda")
        s = torch.cuda.current_stream()

        def run():
            _native.check(lib.hz_knn(0, ctypes.c_void_p(Xd.data_ptr()), n, 50, k,
                                     ctypes.c_void_p(idx.data_ptr()),
                                     ctypes.c_void_p(dist.data_ptr()),
                                     ctypes.c_void_p(ws.data_ptr()), ws.numel(),
                                     ctypes.c_void_p(s.cuda_stream)), "hz_knn")

        run()
        torch.cuda.synchronize()
        reps = 3 if n <= 300_000 else 1
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(s)
        for _ in range(reps):
            run()
        e1.record(s)
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / reps
        tflops = 2.0 * n * n * 50 / (ms * 1e-3) / 1e12
        out.append({"n_cells": n, "n_pcs": 50, "k": k, "ms": round(ms, 3),
                    "roofline": {"bound": "mfma", "achieved": round(tflops, 2), "peak": 157.3,
                                 "unit": "TFLOP/s", "frac": round(tflops / 157.3, 4)}})
        print(f"[bench] knn n={n}: {ms:.2f} ms, {tflops:.1f} TFLOP/s", file=sys.stderr, flush=True)
        del Xd, idx, dist, ws
    print(json.dumps({"metric": "brute-force kNN time (k=15, euclidean, FP32)",
                      "unit": "ms", "higher_is_better": False, "results": out}), flush=True)
    return 0


def end_to_end(args) -> int:
    """The drop-in call with host buffers: ``pp.harmony_integrate`` on an
    AnnData-like object whose ``obsm["X_pca"]`` is a host array.  The timed
    call includes the host-to-device copy, the k-means initialisation, 10
    Harmony iterations of 20 clustering iterations (early stop off) and the
    copy of the result back to the host."""
    import pandas as pd
    import torch

    import rapids_singlecell_amd as rsc

    X, codes, B, K = make_data(args.workload)
    n = X.shape[0]

    class _AD:
        def __init__(self):
            self.obsm = {"X_pca": X}
            self.obs = pd.DataFrame({"batch": pd.Categorical.from_codes(codes, list(range(B)))})

    kw = dict(dtype=np.float32, flavor="harmony1", n_clusters=K, max_iter_clustering=20,
              tol_clustering=float("-inf"), tol_harmony=float("-inf"), sigma=0.1, theta=2.0,
              block_proportion=0.05, correction_method="batched", precision=args.precision)
    rsc.pp.harmony_integrate(_AD(), "batch", **dict(kw, max_iter_harmony=1))  # warm-up
    torch.cuda.synchronize()
    ad = _AD()
    t0 = time.perf_counter()
    rsc.pp.harmony_integrate(ad, "batch", max_iter_harmony=10, **kw)
    wall = time.perf_counter() - t0
    assert isinstance(ad.obsm["X_pca_harmony"], np.ndarray)
    print(json.dumps({
        "metric": METRIC + " (drop-in call, host buffers, PCIe-inclusive)",
        "value": round(n * 10 / wall, 1), "unit": "cells·iters/s", "n_gpus": 1,
        "higher_is_better": True, "wall_s": round(wall, 3), "harmony_iters": 10,
        "workload": args.workload, "precision": args.precision,
        "includes": "H2D copy, k-means init, 10 Harmony iterations, D2H copy",
    }), flush=True)
    return 0


def throughput_leg(X, codes, B, K, precision: str, warmup: int, steps: int,
                   with_roofline: bool = True) -> dict:
    """Time ``steps`` Harmony iterations of one job after ``warmup`` untimed ones
    (inputs resident in HBM), then two more under live HIP-event timing."""
    import torch

    prec = {"native": 0, "split_bf16": 1}[precision]
    job = build_job(X, codes, B, K, prec)
    job.init()
    torch.cuda.synchronize()
    it = 0
    for _ in range(warmup):
        job.step(it)
        it += 1
    torch.cuda.synchronize()
    dt = time_steps(job, it, steps)
    it += steps
    n = X.shape[0]
    leg = {"precision": precision, "value": round(n * steps / dt, 1),
           "ms_per_step": round(1e3 * dt / steps, 3), "steps": steps}
    print(f"[bench] {precision}: {steps} steps in {dt:.3f} s -> {leg['value']:.4g} cells·iters/s",
          file=sys.stderr, flush=True)
    if with_roofline:
        leg["roofline"], leg["kernels"] = roofline(job, it, 2, X.dtype.itemsize, precision)
    del job
    torch.cuda.empty_cache()
    return leg


def _fixed_run(X, codes, B, K, Y0, precision: int, iters: int):
    """``iters`` Harmony iterations from the centroids Y0; returns the final
    Z_corr, R and Y = L2(R^T Z_norm) as FP64 device tensors."""
    import torch

    job = build_job(X, codes, B, K, precision)
    job.init(torch.as_tensor(Y0))
    for i in range(iters):
        job.step(i)
    Z = job.Z_hat.double()
    R = job.R.double()
    Y = R.T @ job.Z_norm.double()
    Y = Y / torch.linalg.norm(Y, dim=1, keepdim=True).clamp_min(1e-12)
    del job
    return Z, R, Y


def error_leg(X, codes, B, K, iters: int = 2) -> dict:
    """Relative Frobenius error of Z_corr, R and Y of the FP32 storage paths
    (plain FP32 and split-BF16) against the FP64 path, all from the same
    initial centroids, after ``iters`` Harmony iterations.  The FP64 path
    equals the FP64 oracle within 1e-8 (tests at 10k and 200k cells)."""
    import torch

    Y0 = initial_centroids(X, K)
    ref = _fixed_run(X.astype(np.float64), codes, B, K, Y0, 0, iters)
    out = {"reference": f"FP64 GPU path, {iters} Harmony iterations x 20 clustering iterations, "
                        "same initial centroids"}
    for name, prec in (("native", 0), ("split_bf16", 1)):
        got = _fixed_run(X, codes, B, K, Y0.astype(np.float32), prec, iters)
        out[name] = {k: float(torch.linalg.norm(g - r) / torch.linalg.norm(r))
                     for k, g, r in zip(("Z_corr", "R", "Y"), got, ref)}
        del got
    del ref
    torch.cuda.empty_cache()
    print(f"[bench] errors vs FP64: {out}", file=sys.stderr, flush=True)
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="c3", choices=sorted(WORKLOADS))
    ap.add_argument("--precision", default="split_bf16", choices=["native", "split_bf16"],
                    help="precision of the --served/--batched/--end-to-end/--pmc-run modes")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-errors", action="store_true")
    ap.add_argument("--no-roofline", action="store_true")
    ap.add_argument("--fp64-steps", type=int, default=2,
                    help="timed steps of the FP64 leg of the default run (0 = skip)")
    ap.add_argument("--batch-copies", type=int, default=4,
                    help="instances of the workload in the batched leg of the default run "
                         "(1 = skip)")
    ap.add_argument("--served-leg", type=float, default=12.0,
                    help="seconds of the served leg of the default run (0 = skip)")
    ap.add_argument("--pmc-run", action="store_true",
                    help="short run for hardware-counter passes: a 256 MiB device copy "
                         "(calibration kernel of known size) then 1 Harmony iteration")
    ap.add_argument("--served", action="store_true",
                    help="served form: Poisson job stream, p50/p99 latency")
    ap.add_argument("--served-rate", type=float, default=20.0)
    ap.add_argument("--served-duration", type=float, default=30.0)
    ap.add_argument("--served-slots", type=int, default=4)
    ap.add_argument("--served-pool", type=int, default=12)
    ap.add_argument("--served-seed", type=int, default=0)
    ap.add_argument("--end-to-end", action="store_true",
                    help="time the drop-in harmony_integrate call from host arrays")
    ap.add_argument("--batched", action="store_true",
                    help="batched form: 64 uneven instances queued at once")
    ap.add_argument("--batched-n", type=int, default=64)
    ap.add_argument("--batched-seed", type=int, default=0)
    ap.add_argument("--multi", action="store_true",
                    help="several covariates (16 batches + 64 donors) at --multi-cells")
    ap.add_argument("--multi-cells", type=int, default=1_300_000)
    ap.add_argument("--knn", action="store_true",
                    help="time the brute-force kNN of the neighbour graph")
    ap.add_argument("--knn-sizes", type=int, nargs="+", default=[200_000, 1_300_000])
    ap.add_argument("--tune", action="append", default=[],
                    help="library tuning knob key=value (row_chunk, centroid_wg, no_sweep, sweep_wg, "
                         "sweep_plain, spin_limit, block_update)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}); the Harmony path "
              "runs on a single MI355X", file=sys.stderr)
        return 2

    if args.served:
        return served(args)
    if args.batched:
        return batched(args)
    if args.end_to_end:
        return end_to_end(args)
    if args.knn:
        return knn_bench(args)
    if args.multi:
        return multi_bench(args)

    import torch

    from rapids_singlecell_amd import _native

    _native.load()
    tune = {k: int(v) for k, v in (t.split("=", 1) for t in args.tune)}
    if tune:
        _native.set_tuning(**tune)
    X, codes, B, K = make_data(args.workload)
    n, d = X.shape
    print(f"[bench] workload {args.workload}: N={n} d={d} K={K} B={B}", file=sys.stderr,
          flush=True)
    if args.pmc_run:
        job = build_job(X, codes, B, K, {"native": 0, "split_bf16": 1}[args.precision])
        job.init()
        torch.cuda.synchronize()
        # calibration: a plain copy reads and writes exactly COPY_BYTES
        src = torch.ones(COPY_BYTES // 4, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
        job.step(0)
        torch.cuda.synchronize()
        print(json.dumps({"pmc_run": True, "copy_bytes": COPY_BYTES, "workload": args.workload,
                          "precision": args.precision}), flush=True)
        return 0

    roof_on = not args.no_roofline
    legs = {p: throughput_leg(X, codes, B, K, p, args.warmup, args.steps, roof_on)
            for p in ("native", "split_bf16")}
    head = legs["split_bf16"]
    # the reference's default dtype is float64: the FP64 path (per-block
    # launches, FP64 kernels) on the same workload, a few steps
    fp64 = None
    if args.fp64_steps > 0:
        fp64 = throughput_leg(X.astype(np.float64), codes, B, K, "native", 1, args.fp64_steps,
                              roof_on)
    batch = None
    if args.batch_copies > 1:
        batch = batched_leg(X, codes, B, K, "split_bf16", args.batch_copies, args.warmup,
                            args.steps)
    errors = None if args.no_errors else error_leg(X, codes, B, K)
    served_out = None
    if args.served_leg > 0:
        print("[bench] served leg ...", file=sys.stderr, flush=True)
        sv = run_served(args, args.served_leg)
        served_out = {k: sv[k] for k in ("value", "latency", "jobs", "offered_rate_jobs_per_s",
                                         "duration_s", "wall_s", "slots", "n_failures")}
    cpu = None
    if not args.no_cpu:
        print("[bench] cpu baseline ...", file=sys.stderr, flush=True)
        cpu = cpu_baseline(X, codes, B, K)
    value_leg = batch if batch is not None else head
    out = {
        "metric": METRIC,
        "value": value_leg["value"],
        "unit": "cells·iters/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": value_leg["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 storage, split-bf16 products, f32 accumulation",
        "data": "synthetic: Gaussian-mixture PCA embedding with per-batch shifts (no datasets offline)",
        "config": {
            "workload": workload_text(args.workload),
            "value_leg": (f"split_bf16, {batch['instances']} instances of the workload stepped "
                          "together (one multi-job clustering call per Harmony iteration)"
                          if batch is not None else "split_bf16, one job"),
            "tuning": _native.get_tuning(),
            "n_cells": n, "n_pcs": d, "n_clusters": K, "n_batches": B,
        },
        "fp32": {k: legs["native"][k] for k in ("value", "ms_per_step")},
        "split_bf16_single": {k: head[k] for k in ("value", "ms_per_step")},
        "fp64": ({k: fp64[k] for k in ("value", "ms_per_step", "roofline", "kernels")}
                 if fp64 else None),
        "split_bf16_batched": batch,
        "value_vs_fp32": round(value_leg["value"] / legs["native"]["value"], 3),
        "split_bf16_single_vs_fp32": round(head["value"] / legs["native"]["value"], 3),
        "rel_err_vs_fp64": errors,
        "latency": served_out["latency"] if served_out else None,
        "served": served_out,
        "roofline": head.get("roofline"),
        "roofline_fp32": legs["native"].get("roofline"),
        "cpu_baseline": cpu,
        "kernels": head.get("kernels"),
        "kernels_fp32": legs["native"].get("kernels"),
    }
    if cpu:
        out["fp32_vs_cpu"] = round(legs["native"]["value"] / cpu["value"], 1)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
