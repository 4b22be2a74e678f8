"""Batched Harmony: several independent instances advanced together.

``harmonize_many`` runs many integration problems at once on one stream.
Per Harmony iteration, the instances whose kernels share a shape (precision,
cluster-count class, PC-count class) run their clustering step through ONE
call of ``hz_clustering_loop_batched``: every clustering iteration does each
instance's centroid pass and then a single cooperative launch in which every
instance's block sweep runs on its own share of the device's workgroups.  The
ridge corrections then run per instance.  Instances that cannot use the
block sweep (float64, several covariates, very small blocks) run their
clustering step alone, in the same loop.

Each instance's result is the one ``harmonize`` gives for it alone: the block
sums and objective terms are fixed-point, so the share of workgroups an
instance gets does not change a bit of its state.
"""

from __future__ import annotations

import ctypes
import warnings

import numpy as np
import torch

from ... import _native
from ._helper import _get_batch_codes, _get_theta_array
from ._job import HarmonyJob

_U32 = 0xFFFFFFFF


class HzJob(ctypes.Structure):
    """``hz_job`` of include/harmony_hip.h."""

    _fields_ = [
        ("dims", _native.HzDims),
        ("Z_norm", ctypes.c_void_p),
        ("R", ctypes.c_void_p),
        ("E", ctypes.c_void_p),
        ("O", ctypes.c_void_p),
        ("Pr_b", ctypes.c_void_p),
        ("cats", ctypes.c_void_p),
        ("theta", ctypes.c_void_p),
        ("sigma", ctypes.c_double),
        ("tol", ctypes.c_double),
        ("max_iter", ctypes.c_int),
        ("seed", ctypes.c_uint32),
        ("stabilized", ctypes.c_int),
        ("last_obj", ctypes.c_void_p),
        ("n_iter", ctypes.c_int),
        ("ws", ctypes.c_void_p),
        ("ws_bytes", ctypes.c_size_t),
    ]


def _nt_class(K: int) -> int:
    nt = -(-K // 16)
    return 4 if nt <= 4 else 7 if nt <= 7 else 8 if nt <= 8 else 16


def _sm_class(d: int, precision: int) -> int:
    if precision == 1:
        return -(-d // 32)
    s = -(-d // 4)
    for c in (4, 8, 13, 16, 25):
        if s <= c:
            return c
    return 32


def sweep_class(job: HarmonyJob):
    """Kernel shape of a job's block sweep, or None when the job cannot run
    in the batched sweep (``hz_batched_eligible``: float32, one covariate,
    at most 256 clusters and 64 blocks, and a sweep whose LDS fits)."""
    cached = getattr(job, "_sweep_class", False)
    if cached is not False:
        return cached
    c = None
    if job.dtype == torch.float32 and job.n_covariates == 1:
        ok = ctypes.c_int(0)
        _native.check(_native.load().hz_batched_eligible(ctypes.byref(job.dims), ctypes.byref(ok)),
                      "hz_batched_eligible")
        if ok.value:
            c = (job.precision, _nt_class(job.n_clusters), _sm_class(job.n_pcs, job.precision),
                 job.max_iter_clustering)
    job._sweep_class = c
    return c


_bws_cache: dict = {}


def _batched_workspace(n: int, max_iter: int, device, stream) -> torch.Tensor:
    lib = _native.load()
    nbytes = lib.hz_batched_workspace_bytes(n, max_iter)
    key = (device, stream.cuda_stream)
    buf = _bws_cache.get(key)
    if buf is None or buf.numel() < nbytes:
        with torch.cuda.stream(stream):
            buf = torch.empty(max(nbytes, 1 << 16), dtype=torch.uint8, device=device)
        _bws_cache[key] = buf
    return buf


def clustering_batched(jobs: list[HarmonyJob], harmony_iter: int) -> list[int]:
    """One Harmony iteration's clustering step of several jobs of one sweep
    class, in one library call; returns each job's clustering iterations."""
    lib = _native.load()
    n = len(jobs)
    stream = jobs[0].stream
    arr = (HzJob * n)()
    for i, job in enumerate(jobs):
        if job.stream.cuda_stream != stream.cuda_stream:
            raise ValueError("batched jobs must share one stream")
        h = arr[i]
        h.dims = job.dims
        h.Z_norm = job.Z_norm.data_ptr()
        h.R = job.R.data_ptr()
        h.E = job.E.data_ptr()
        h.O = job.O.data_ptr()
        h.Pr_b = job.Pr_b.data_ptr()
        h.cats = job.cats.data_ptr()
        h.theta = job.theta.data_ptr()
        h.sigma = job.sigma
        h.tol = job.tol_clustering
        h.max_iter = job.max_iter_clustering
        h.seed = (job.random_state + harmony_iter * 1000003) & _U32
        h.stabilized = job.stabilized
        h.last_obj = job.last_obj.data_ptr()
        h.ws = job.ws.data_ptr()
        h.ws_bytes = job.ws_bytes
    bws = _batched_workspace(n, jobs[0].max_iter_clustering, jobs[0].dev, stream)
    rc = lib.hz_clustering_loop_batched(ctypes.cast(arr, ctypes.c_void_p), n,
                                        ctypes.c_void_p(bws.data_ptr()), bws.numel(),
                                        ctypes.c_void_p(stream.cuda_stream))
    _native.check(rc, "hz_clustering_loop_batched")
    return [arr[i].n_iter for i in range(n)]


def clustering_step(jobs: list[HarmonyJob], harmony_iter: int,
                    max_group: int | None = None) -> list[int]:
    """Clustering step of Harmony iteration ``harmony_iter`` for every job:
    jobs of one sweep class together (at most ``max_group`` per call), the
    others alone.  Returns each job's clustering iterations."""
    out = [0] * len(jobs)
    groups: dict = {}
    for i, job in enumerate(jobs):
        c = sweep_class(job)
        if c is None:
            out[i] = job.clustering(harmony_iter)
        else:
            groups.setdefault(c, []).append(i)
    for members in groups.values():
        step = max_group or len(members)
        for g0 in range(0, len(members), step):
            part = members[g0:g0 + step]
            if len(part) == 1:
                out[part[0]] = jobs[part[0]].clustering(harmony_iter)
                continue
            for i, nc in zip(part, clustering_batched([jobs[i] for i in part], harmony_iter)):
                out[i] = nc
    return out


def step_many(jobs: list[HarmonyJob], harmony_iter: int, max_group: int | None = None) -> list[int]:
    """One Harmony iteration (clustering, then each job's correction) of
    several jobs that share a stream."""
    out = clustering_step(jobs, harmony_iter, max_group)
    for job in jobs:
        job.correction()
    return out


def harmonize_many(problems: list[dict], *, max_group: int | None = None,
                   stream: torch.cuda.Stream | None = None, record: list | None = None):
    """Run several Harmony problems together.

    Each problem is a dict with ``X`` (device tensor, n x d), ``obs`` (batch
    table) and ``key`` plus ``harmonize`` keyword arguments (``n_clusters``,
    ``theta``, ``sigma``, ``max_iter_harmony``, ``max_iter_clustering``,
    ``tol_harmony``, ``tol_clustering``, ``ridge_lambda``,
    ``stabilized_penalty``, ``dynamic_lambda``, ``alpha``,
    ``batch_prune_threshold``, ``correction_method``, ``block_proportion``,
    ``random_state``, ``precision``, ``init_centroids``).  ``max_group``
    caps the instances per batched call.  Returns the Z_hat tensors in
    problem order; ``record`` (a list) receives one dict per problem with the
    objectives and clustering iteration counts.
    """
    if not problems:
        return []
    dev = problems[0]["X"].device
    if stream is None:
        stream = torch.cuda.current_stream(dev)
    precisions = {"native": 0, "fp32": 0, "fp64": 0, "split_bf16": 1}
    jobs, meta = [], []
    with torch.cuda.stream(stream):
        for p in problems:
            X = p["X"]
            if not X.is_cuda:
                raise TypeError("harmonize_many needs device tensors")
            np_dtype = np.float32 if X.dtype == torch.float32 else np.float64
            n = X.shape[0]
            codes = _get_batch_codes(p["obs"], p.get("key", "batch"))
            B = codes.cat.categories.size
            N_b = np.asarray(codes.value_counts(sort=False).values, dtype=np_dtype)
            cats = np.asarray(codes.cat.codes.values, dtype=np.int32)
            K = p.get("n_clusters") or max(int(min(100, n / 30)), 2)
            theta = _get_theta_array(p.get("theta", 2.0), B, np_dtype)
            tau = p.get("tau", 0)
            if tau > 0:
                theta = theta * (1 - np.exp(-N_b / (K * tau)) ** 2)
            job = HarmonyJob(
                X, cats, B, n_clusters=K, theta_np=np.ascontiguousarray(theta.astype(np_dtype)),
                N_b_np=N_b, sigma=p.get("sigma", 0.1),
                block_proportion=p.get("block_proportion", 0.05),
                max_iter_clustering=p.get("max_iter_clustering", 200),
                tol_clustering=p.get("tol_clustering", 1e-5),
                ridge_lambda=p.get("ridge_lambda", 1.0),
                stabilized_penalty=p.get("stabilized_penalty", True),
                dynamic_lambda=p.get("dynamic_lambda", True), alpha=p.get("alpha", 0.2),
                batch_prune_threshold=p.get("batch_prune_threshold", 1e-5),
                correction_method=p.get("correction_method") or "batched",
                random_state=p.get("random_state", 0),
                precision=precisions[p.get("precision", "native")], stream=stream)
            Y0 = p.get("init_centroids")
            if Y0 is not None:
                Y0 = torch.as_tensor(np.asarray(Y0), dtype=X.dtype, device=dev)
            job.init(Y0)
            jobs.append(job)
            meta.append({"max_iter": int(p.get("max_iter_harmony", 10)),
                         "tol": float(p.get("tol_harmony", 1e-4)), "objectives": [],
                         "clustering_iters": [], "done": False})
        objs0 = torch.stack([j.obj.double().reshape(()) for j in jobs]).cpu().numpy()
    for m, o in zip(meta, objs0):
        m["objectives"].append(float(o))

    it = 0
    while True:
        active = [i for i, m in enumerate(meta) if not m["done"] and it < m["max_iter"]]
        if not active:
            break
        for i, nc in zip(active, clustering_step([jobs[i] for i in active], it, max_group)):
            meta[i]["clustering_iters"].append(nc)
        with torch.cuda.stream(stream):
            lasts = torch.stack([jobs[i].last_obj.double().reshape(()) for i in active]).cpu()
        for i, v in zip(active, lasts.numpy()):
            meta[i]["objectives"].append(float(v))
            jobs[i].correction()
            objs = meta[i]["objectives"]
            if (objs[-2] - objs[-1]) < meta[i]["tol"] * abs(objs[-2]):
                meta[i]["done"] = True
        it += 1
    stream.synchronize()
    for m in meta:
        if not m["done"] and len(m["clustering_iters"]) >= m["max_iter"] and np.isfinite(m["tol"]):
            warnings.warn("Harmony did not converge. Consider increasing the number of iterations")
            break
    if record is not None:
        record.extend({"objectives": m["objectives"], "clustering_iters": m["clustering_iters"],
                       "R": j.R, "O": j.O, "E": j.E} for m, j in zip(meta, jobs))
    return [j.Z_hat for j in jobs]
