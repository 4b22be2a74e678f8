"""``pp.neighbors`` with the brute-force search on the GPU: the consumer of
the integrated embedding (``obsm["X_pca_harmony"]``).

Follows preprocessing/_neighbors/__init__.py:35-255 of rapids_singlecell:
same signature, representation choice, self-distance fix, result keys and
``uns`` record.  The exact k nearest neighbours come from ``hz_knn``
(matrix-core distance tiles with a fused top-k); the UMAP membership weights
from ``hz_umap_weights``.  As in the reference, the distance and connectivity
matrices are assembled as SciPy CSR on the host; the UMAP fuzzy union
(W + W^T - W o W^T), the Gaussian kernel and the Jaccard index are sparse
operations on that host graph.
"""

from __future__ import annotations

import copy as _copy
import ctypes
from types import MappingProxyType

import numpy as np
import torch
from scipy import sparse

from .. import _native

_METRICS = {"euclidean": 0, "l2": 0, "sqeuclidean": 1, "cosine": 2}
_ALGORITHMS = ("brute", "cagra", "ivfflat", "ivfpq", "nn_descent", "all_neighbors", "mg_ivfflat",
               "mg_ivfpq")


def _choose_representation(adata, use_rep=None, n_pcs=None):
    """tools/_utils.py semantics: ``use_rep`` names an ``obsm`` key or "X";
    without it, ``X_pca`` when present (or ``.X`` for fewer than 50
    variables); ``n_pcs`` keeps the first columns."""
    if use_rep is None:
        if "X_pca" in adata.obsm:
            X = adata.obsm["X_pca"]
        elif getattr(adata, "X", None) is not None:
            X = adata.X
        else:
            raise ValueError("no representation: pass use_rep or set obsm['X_pca']")
    elif use_rep == "X":
        X = adata.X
    else:
        if use_rep not in adata.obsm:
            raise KeyError(f"Did not find {use_rep!r} in .obsm")
        X = adata.obsm[use_rep]
    if n_pcs is not None:
        if n_pcs > X.shape[1]:
            raise ValueError(f"{X.shape[1]} columns available, n_pcs={n_pcs} requested")
        X = X[:, :n_pcs]
    return X


def knn(X, k: int, metric: str = "euclidean"):
    """Exact kNN of every row among all rows (device tensors ``indices`` int32
    and ``distances`` float32, n x k, nearest first)."""
    if metric not in _METRICS:
        raise ValueError(f"metric must be one of {sorted(_METRICS)} for the brute-force search")
    lib = _native.load()
    Xd = torch.as_tensor(np.ascontiguousarray(X) if not isinstance(X, torch.Tensor) else X)
    Xd = Xd.to(device="cuda", dtype=torch.float32).contiguous()
    n, d = Xd.shape
    m = _METRICS[metric]
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    idx = torch.empty((n, k), dtype=torch.int32, device=Xd.device)
    dist = torch.empty((n, k), dtype=torch.float32, device=Xd.device)
    nb = lib.hz_knn_workspace_bytes(n, d, m)
    ws = torch.empty(max(nb, 256), dtype=torch.uint8, device=Xd.device)
    _native.check(lib.hz_knn(m, ctypes.c_void_p(Xd.data_ptr()), n, d, k,
                             ctypes.c_void_p(idx.data_ptr()), ctypes.c_void_p(dist.data_ptr()),
                             ctypes.c_void_p(ws.data_ptr()), ws.numel(), s), "hz_knn")
    return idx, dist


def umap_weights(idx: torch.Tensor, dist: torch.Tensor):
    """UMAP membership weights of the directed kNN graph (device, n x k)."""
    lib = _native.load()
    n, k = idx.shape
    w = torch.empty_like(dist)
    mean = float(dist.double().mean().item())
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    _native.check(lib.hz_umap_weights(ctypes.c_void_p(dist.data_ptr()),
                                      ctypes.c_void_p(idx.data_ptr()), n, k, mean,
                                      ctypes.c_void_p(w.data_ptr()), None, None, s),
                  "hz_umap_weights")
    return w


def _connectivities(idx_h, dist_h, w_h, n, n_neighbors, method):
    k = idx_h.shape[1]
    if method == "umap":
        W = sparse.csr_matrix((w_h.ravel(), (np.repeat(np.arange(n), k), idx_h.ravel())),
                              shape=(n, n))
        W.eliminate_zeros()
        T = W.T.tocsr()
        out = sparse.csr_matrix(W + T - W.multiply(T))
        out.eliminate_zeros()
        return out.astype(np.float32)
    if method == "gauss":
        d_sq = dist_h.astype(np.float64) ** 2
        sig_sq = np.median(d_sq[:, 1:], axis=1)
        sig = np.sqrt(sig_sq)
        rows = np.repeat(np.arange(n), k - 1)
        cols = idx_h[:, 1:].ravel()
        den = sig_sq[rows] + sig_sq[cols]
        vals = np.sqrt(2.0 * sig[rows] * sig[cols] / den) * np.exp(-d_sq[:, 1:].ravel() / den)
        W = sparse.coo_matrix((vals, (rows, cols)), shape=(n, n)).tocsr()
        return W.maximum(W.T.tocsr()).tocsr().astype(np.float32)
    if method == "jaccard":
        kk = n_neighbors - 1
        rows = np.repeat(np.arange(n), kk)
        cols = idx_h[:, 1:].ravel()
        A = sparse.csr_matrix((np.ones(n * kk, np.float32), (rows, cols)), shape=(n, n))
        shared = np.asarray(A[rows].multiply(A[cols]).sum(axis=1)).ravel()
        jac = shared / (2 * kk - shared)
        m = jac != 0
        W = sparse.coo_matrix((jac[m], (rows[m], cols[m])), shape=(n, n)).tocsr()
        return ((W + W.T) / 2).tocsr().astype(np.float32)
    raise ValueError(f"method must be 'umap', 'gauss' or 'jaccard', got {method!r}")


def neighbors(adata, n_neighbors: int = 15, n_pcs: int | None = None, *, use_rep: str | None = None,
              random_state=0, algorithm: str = "brute", metric: str = "euclidean",
              metric_kwds=MappingProxyType({}), algorithm_kwds=MappingProxyType({}),
              method: str = "umap", key_added: str | None = None, copy: bool = False):
    """Neighbourhood graph of the observations (see module docstring).

    Parameters and results are those of ``rapids_singlecell.pp.neighbors``;
    the search is exact (``algorithm="brute"``; the approximate cuVS indexes
    are not provided) with the ``euclidean``, ``sqeuclidean`` and ``cosine``
    metrics and ``n_neighbors <= 32``.
    """
    adata = _copy.deepcopy(adata) if copy else adata
    if algorithm not in _ALGORITHMS:
        raise ValueError(f"Invalid algorithm '{algorithm}' for KNN. Valid options are: "
                         f"{_ALGORITHMS}.")
    if algorithm != "brute":
        raise NotImplementedError(f"algorithm={algorithm!r}: only the exact 'brute' search is "
                                  "provided")
    X = _choose_representation(adata, use_rep=use_rep, n_pcs=n_pcs)
    n = X.shape[0]
    idx, dist = knn(X, n_neighbors, metric)
    dist[:, 0] = 0.0  # _fix_self_distances
    w = umap_weights(idx, dist) if method == "umap" else None
    idx_h = idx.cpu().numpy().astype(np.int64)
    dist_h = dist.cpu().numpy()
    w_h = w.cpu().numpy() if w is not None else None
    k = n_neighbors
    distances = sparse.csr_matrix((dist_h.ravel(), idx_h.ravel(), np.arange(0, n * k + 1, k)),
                                  shape=(n, n))
    connectivities = _connectivities(idx_h, dist_h, w_h, n, n_neighbors, method)
    params = dict(n_neighbors=n_neighbors, method="rapids", random_state=random_state,
                  metric=metric,
                  **({"metric_kwds": dict(metric_kwds)} if metric_kwds else {}),
                  **({"algorithm_kwds": dict(algorithm_kwds)} if algorithm_kwds else {}),
                  **({"use_rep": use_rep} if use_rep is not None else {}),
                  **({"n_pcs": n_pcs} if n_pcs is not None else {}))
    key = key_added if key_added is not None else "neighbors"
    conns_key = "connectivities" if key == "neighbors" else key + "_connectivities"
    dists_key = "distances" if key == "neighbors" else key + "_distances"
    if not hasattr(adata, "uns"):
        adata.uns = {}
    if not hasattr(adata, "obsp"):
        adata.obsp = {}
    adata.uns[key] = {"connectivities_key": conns_key, "distances_key": dists_key,
                      "params": params}
    adata.obsp[dists_key] = distances
    adata.obsp[conns_key] = connectivities
    return adata if copy else None
