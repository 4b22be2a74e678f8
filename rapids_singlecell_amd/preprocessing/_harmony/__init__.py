"""Harmony driver over the HIP library.

Restates ``harmonize`` of the reference (preprocessing/_harmony/__init__.py:
41-289): validation, batch bookkeeping, initialisation, the Harmony outer loop
with clustering + correction, and the convergence test.  All state (Z_norm, R,
E, O, Y, Z_hat) stays in device memory; per Harmony iteration the host reads
one scalar (the clustering objective) and, during clustering, a done flag
every few iterations.
"""

from __future__ import annotations

import ctypes
import warnings

import numpy as np
import torch

from ... import _native
from ._helper import _get_batch_codes, _get_covariates, _get_theta_array, _get_theta_levels
from ._job import HarmonyJob

_SUPPRESS_PENALTY = 1e30
COLSUM_ALGOS = ["columns", "atomics", "gemm", "benchmark", None]
_U32 = 0xFFFFFFFF
_PRECISIONS = {"native": 0, "fp32": 0, "fp64": 0, "split_bf16": 1}


def _stream_handle(device: torch.device) -> ctypes.c_void_p:
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)


def _ptr(t: torch.Tensor) -> ctypes.c_void_p:
    return ctypes.c_void_p(t.data_ptr())


def _to_device(x, device, dtype):
    if isinstance(x, torch.Tensor):
        return x.to(device=device, dtype=dtype).contiguous()
    return torch.as_tensor(np.ascontiguousarray(x), dtype=dtype, device=device)


def kmeans_centroids(Z_norm: torch.Tensor, n_clusters: int, random_state: int,
                     max_iter: int = 25) -> torch.Tensor:
    """Initial centroids from the library's seeded Lloyd k-means on the GPU."""
    if not Z_norm.is_cuda:
        raise TypeError("kmeans_centroids needs a device tensor")
    lib = _native.load()
    n, d = Z_norm.shape
    dsz = Z_norm.element_size()
    dev = Z_norm.device
    nbytes = lib.hz_kmeans_workspace_bytes(dsz, n, d, n_clusters)
    if nbytes == 0:
        _native.check(-4, "hz_kmeans_workspace_bytes")
    ws = torch.empty(nbytes, dtype=torch.uint8, device=dev)
    Y = torch.empty((n_clusters, d), dtype=Z_norm.dtype, device=dev)
    rc = lib.hz_kmeans(dsz, _ptr(Z_norm), n, d, n_clusters, random_state & _U32, max_iter,
                       _ptr(Y), _ptr(ws), nbytes, _stream_handle(dev))
    _native.check(rc, "hz_kmeans")
    return Y


def l2_normalize(X: torch.Tensor) -> torch.Tensor:
    if not X.is_cuda:
        raise TypeError("l2_normalize needs a device tensor")
    X = X.contiguous()
    lib = _native.load()
    out = torch.empty_like(X)
    rc = lib.hz_l2_normalize(X.element_size(), _ptr(X), _ptr(out), X.shape[0], X.shape[1],
                             _stream_handle(X.device))
    _native.check(rc, "hz_l2_normalize")
    return out


def harmonize(
    Z,
    batch_mat,
    batch_key,
    *,
    n_clusters: int | None = None,
    max_iter_harmony: int = 10,
    max_iter_clustering: int = 200,
    tol_harmony: float = 1e-4,
    tol_clustering: float = 1e-5,
    ridge_lambda: float = 1.0,
    sigma: float = 0.1,
    block_proportion: float = 0.05,
    theta=2.0,
    tau: int = 0,
    correction_method: str | None = None,
    colsum_algo: str | None = None,
    random_state: int = 0,
    stabilized_penalty: bool = True,
    dynamic_lambda: bool = True,
    alpha: float = 0.2,
    batch_prune_threshold: float | None = 1e-5,
    verbose: bool = False,
    init_centroids=None,
    precision: str = "native",
    record: dict | None = None,
    multi_covariate: bool = False,
) -> torch.Tensor:
    """Integrate data using the Harmony algorithm (device tensors in and out).

    Parameters and errors follow the reference ``harmonize``.  Extensions:
    ``init_centroids`` (K x d) replaces the k-means initialiser (parity runs
    inject the same centroids into the oracle); ``precision`` selects the GEMM
    arithmetic; ``record`` receives objectives and clustering iteration counts;
    ``multi_covariate=True`` with several keys treats each key as a covariate
    of its own (stacked one-hot Phi: O and E per level, a cell's penalty is
    the product of its levels' penalties, the ridge correction solves the
    full normal equations) instead of joining the keys into one categorical.
    ``theta`` is then a scalar, one value per key or one value per level.
    ``colsum_algo`` is validated and otherwise ignored: column sums are fused
    into the row kernel.
    """
    if not isinstance(Z, torch.Tensor):
        raise TypeError("harmonize expects a device tensor; use harmony_integrate for AnnData")
    if Z.dtype not in (torch.float32, torch.float64):
        raise TypeError(f"unsupported dtype {Z.dtype}")
    dev = Z.device
    Z = Z.contiguous()
    n_cells, n_pcs = Z.shape
    dsz = Z.element_size()
    np_dtype = np.float32 if Z.dtype == torch.float32 else np.float64
    if precision not in _PRECISIONS:
        raise ValueError(f"precision must be one of {sorted(_PRECISIONS)}, got {precision!r}")
    prec = _PRECISIONS[precision]
    if prec == 1 and Z.dtype != torch.float32:
        raise ValueError("precision='split_bf16' needs float32 data")

    multi = bool(multi_covariate) and not isinstance(batch_key, str) and len(batch_key) >= 2
    group_levels = None
    n_levels0 = 0
    if multi:
        # one covariate per key: cells are grouped by their tuple of levels
        _, n_levels, group_levels, cats_np = _get_covariates(batch_mat, list(batch_key))
        n_batches = int(group_levels.shape[0])
        N_b_np = np.bincount(group_levels[cats_np].ravel(),
                             minlength=int(sum(n_levels))).astype(np_dtype)
        n_levels0 = int(n_levels[0])
    else:
        # batch bookkeeping (_harmony/__init__.py:146-152)
        batch_codes = _get_batch_codes(batch_mat, batch_key)
        n_batches = batch_codes.cat.categories.size
        N_b_np = np.asarray(batch_codes.value_counts(sort=False).values, dtype=np_dtype)
        cats_np = np.asarray(batch_codes.cat.codes.values, dtype=np.int32)
        if n_cells and cats_np.min() < 0:
            raise ValueError(f"batch key {batch_key!r} has missing values")

    # parameters (_harmony/__init__.py:155-200)
    if max_iter_harmony < 1:
        raise ValueError("max_iter_harmony must be >= 1")
    if n_clusters is None:
        n_clusters = int(min(100, n_cells / 30))
        n_clusters = max(n_clusters, 2)
    assert colsum_algo in COLSUM_ALGOS
    if multi:
        theta_np = _get_theta_levels(theta, n_levels, np_dtype)
    else:
        theta_np = _get_theta_array(theta, n_batches, np_dtype)
    if tau > 0:
        theta_np = theta_np * (1 - np.exp(-N_b_np / (n_clusters * tau)) ** 2)
    theta_np = np.ascontiguousarray(theta_np.astype(np_dtype).ravel())
    assert block_proportion > 0 and block_proportion <= 1
    if dynamic_lambda:
        if not np.isfinite(alpha) or alpha <= 0:
            raise ValueError(
                f"alpha must be a finite positive number when dynamic_lambda=True, got {alpha}."
            )
        if batch_prune_threshold is not None and not (0 <= batch_prune_threshold <= 1):
            raise ValueError(
                f"batch_prune_threshold must be in [0, 1] or None, got {batch_prune_threshold}."
            )
    if correction_method is not None and correction_method not in {"fast", "original", "batched"}:
        raise ValueError("correction_method must be 'fast', 'original', or 'batched'.")
    if correction_method is None:
        nb1 = n_batches + 1
        inv_bytes = n_clusters * nb1 * nb1 * dsz
        correction_method = "batched" if inv_bytes <= (1 << 30) else "fast"
    block_size = int(n_cells * block_proportion)
    if block_size < 1:
        raise ValueError(
            f"block_proportion={block_proportion} gives an empty block for {n_cells} cells"
        )

    if not Z.is_cuda:
        raise TypeError("harmonize needs Z on a HIP device (torch 'cuda' tensor)")
    job = HarmonyJob(
        Z, cats_np, n_batches, n_clusters=n_clusters, theta_np=theta_np, N_b_np=N_b_np,
        sigma=sigma, block_proportion=block_proportion, max_iter_clustering=max_iter_clustering,
        tol_clustering=tol_clustering, ridge_lambda=ridge_lambda,
        stabilized_penalty=stabilized_penalty, dynamic_lambda=dynamic_lambda, alpha=alpha,
        batch_prune_threshold=batch_prune_threshold, correction_method=correction_method,
        random_state=random_state, precision=prec, group_levels=group_levels,
        n_levels0=n_levels0)

    # initialisation (_harmony/__init__.py:206-217, 292-353)
    Y0 = None
    if init_centroids is not None:
        Y0 = _to_device(init_centroids, dev, Z.dtype)
        if tuple(Y0.shape) != (n_clusters, n_pcs):
            raise ValueError(f"init_centroids must have shape {(n_clusters, n_pcs)}")
    job.init(Y0)
    objectives_harmony = [float(job.obj.item())]

    clustering_iters = []
    is_converged = False
    for i in range(max_iter_harmony):
        clustering_iters.append(job.clustering(i))
        objectives_harmony.append(float(job.last_obj.item()))
        job.correction()
        if _is_convergent_harmony(objectives_harmony, tol=tol_harmony):
            is_converged = True
            if verbose:
                print(f"Harmony converged in {i + 1} iterations")
            break
    if not is_converged:
        warnings.warn("Harmony did not converge. Consider increasing the number of iterations")
    if record is not None:
        record.update(objectives=objectives_harmony, clustering_iters=clustering_iters, R=job.R,
                      E=job.E, O=job.O, Z_norm=job.Z_norm, lambda_kb=job.lam, Y0=job.Y0,
                      n_clusters=n_clusters, correction_method=correction_method,
                      group_levels=group_levels)
    return job.Z_hat


def _is_convergent_harmony(objectives_harmony: list, tol: float) -> bool:
    """_harmony/__init__.py:722-734."""
    if len(objectives_harmony) < 2:
        return False
    obj_old = objectives_harmony[-2]
    obj_new = objectives_harmony[-1]
    return (obj_old - obj_new) < tol * np.abs(obj_old)
