"""CPU oracle for the neighbour graph -- TEST INFRASTRUCTURE ONLY.

NumPy / SciPy restatement of what ``rsc.pp.neighbors(algorithm="brute")``
computes on the integrated embedding (preprocessing/_neighbors/__init__.py:
35-255): exact k nearest neighbours, zero self-distances, the sparse distance
matrix and the connectivities of the three methods.  Only ``tests/`` may
import it; the product path never does.

The reference delegates the neighbour search and the UMAP connectivities to
cuML (absent here).  The search is pinned against scikit-learn's brute-force
NearestNeighbors; the UMAP part restates umap-learn's published
``smooth_knn_dist`` / ``compute_membership_strengths`` / fuzzy union (the
algorithm cuML's ``fuzzy_simplicial_set`` implements) and is pinned by its
defining equation and invariants in tests/test_neighbors_oracle.py.  Parity of
the UMAP weights with cuML itself is unpinned.
"""

from __future__ import annotations

import numpy as np
from scipy import sparse


def knn_brute(X, k: int, metric: str = "euclidean", block: int = 2048):
    """Exact kNN in FP64 of every row among all rows, ordered by (distance,
    index).  euclidean / sqeuclidean: direct differences; cosine: 1 - x.y of
    unit rows."""
    X = np.asarray(X, dtype=np.float64)
    n = X.shape[0]
    if metric == "cosine":
        nrm = np.linalg.norm(X, axis=1, keepdims=True)
        X = X / np.where(nrm > 0, nrm, 1.0)
    idx = np.empty((n, k), dtype=np.int64)
    dist = np.empty((n, k), dtype=np.float64)
    sq = np.einsum("ij,ij->i", X, X)
    for r0 in range(0, n, block):
        Q = X[r0:r0 + block]
        if metric == "cosine":
            D = 1.0 - Q @ X.T
        else:
            D = np.maximum(sq[r0:r0 + block, None] + sq[None, :] - 2.0 * (Q @ X.T), 0.0)
        order = np.lexsort((np.broadcast_to(np.arange(n), D.shape), D), axis=1)[:, :k]
        idx[r0:r0 + block] = order
        dd = np.take_along_axis(D, order, axis=1)
        dist[r0:r0 + block] = np.sqrt(dd) if metric == "euclidean" else dd
    return idx, dist


def fix_self_distances(dist, metric):
    """_neighbors/_helper/__init__.py:114-130: column 0 is the self distance."""
    dist = dist.copy()
    if metric != "inner_product":
        dist[:, 0] = 0.0
    return dist


def sparse_distances(idx, dist, n):
    """_neighbors/_neighbors.py:98-131: CSR with k entries per row."""
    k = idx.shape[1]
    return sparse.csr_matrix((dist.ravel(), idx.ravel(), np.arange(0, n * k + 1, k)),
                             shape=(n, n))


def smooth_knn_dist(distances, k, n_iter=64, local_connectivity=1.0, bandwidth=1.0):
    """umap-learn's smooth_knn_dist (FP64 here): rho = the local_connectivity-th
    non-zero distance, sigma by bisection so that sum_{j>=1}
    exp(-max(0, d_j - rho)/sigma) = log2(k) * bandwidth, floored at 1e-3 of
    the row mean (global mean when rho = 0)."""
    distances = np.asarray(distances, dtype=np.float64)
    target = np.log2(k) * bandwidth
    n = distances.shape[0]
    rho = np.zeros(n)
    result = np.zeros(n)
    mean_distances = distances.mean()
    for i in range(n):
        lo, hi, mid = 0.0, np.inf, 1.0
        ith = distances[i]
        nz = ith[ith > 0.0]
        if nz.shape[0] >= local_connectivity:
            index = int(np.floor(local_connectivity))
            interpolation = local_connectivity - index
            if index > 0:
                rho[i] = nz[index - 1]
                if interpolation > 1e-5:
                    rho[i] += interpolation * (nz[index] - nz[index - 1])
            else:
                rho[i] = interpolation * nz[0]
        elif nz.shape[0] > 0:
            rho[i] = nz.max()
        for _ in range(n_iter):
            d = ith[1:] - rho[i]
            psum = np.sum(np.where(d > 0, np.exp(-(np.maximum(d, 0) / mid)), 1.0))
            if abs(psum - target) < 1e-5:
                break
            if psum > target:
                hi = mid
                mid = (lo + hi) / 2.0
            else:
                lo = mid
                mid = mid * 2 if hi == np.inf else (lo + hi) / 2.0
        result[i] = mid
        floor = 1e-3 * (ith.mean() if rho[i] > 0.0 else mean_distances)
        if result[i] < floor:
            result[i] = floor
    return result, rho


def membership_strengths(knn_indices, knn_dists, sigmas, rhos):
    """umap-learn's compute_membership_strengths: 0 for the row itself,
    1 when d <= rho or sigma = 0, exp(-(d - rho)/sigma) otherwise."""
    n, k = knn_indices.shape
    d = knn_dists - rhos[:, None]
    with np.errstate(divide="ignore", invalid="ignore"):
        vals = np.where((d <= 0) | (sigmas[:, None] == 0), 1.0,
                        np.exp(-(d / sigmas[:, None])))
    vals[knn_indices == np.arange(n)[:, None]] = 0.0
    return vals


def fuzzy_union(knn_indices, vals, n, set_op_mix_ratio=1.0):
    """W + W^T - W o W^T (set_op_mix_ratio = 1) of the directed kNN graph."""
    k = knn_indices.shape[1]
    W = sparse.csr_matrix((vals.ravel(), (np.repeat(np.arange(n), k), knn_indices.ravel())),
                          shape=(n, n))
    W.eliminate_zeros()
    T = W.T.tocsr()
    P = W.multiply(T)
    out = set_op_mix_ratio * (W + T - P) + (1.0 - set_op_mix_ratio) * P
    out = sparse.csr_matrix(out)
    out.eliminate_zeros()
    return out


def umap_connectivities(knn_indices, knn_dists, n, k):
    sig, rho = smooth_knn_dist(knn_dists, float(k))
    return fuzzy_union(knn_indices, membership_strengths(knn_indices, knn_dists, sig, rho), n)


def gauss_connectivities(knn_indices, knn_dist, n):
    """_neighbors/_neighbors.py:162-204: sigma_i^2 = median of the squared
    distances of neighbours 1..k-1; w_ij = sqrt(2 s_i s_j / (s_i^2 + s_j^2))
    exp(-d_ij^2 / (s_i^2 + s_j^2)); symmetrised as max(W, W^T)."""
    k = knn_indices.shape[1]
    d_sq = knn_dist ** 2
    sig_sq = np.median(d_sq[:, 1:], axis=1)
    sig = np.sqrt(sig_sq)
    rows = np.repeat(np.arange(n), k - 1)
    cols = knn_indices[:, 1:].ravel()
    den = sig_sq[rows] + sig_sq[cols]
    vals = np.sqrt(2.0 * sig[rows] * sig[cols] / den) * np.exp(-d_sq[:, 1:].ravel() / den)
    W = sparse.coo_matrix((vals, (rows, cols)), shape=(n, n)).tocsr()
    return W.maximum(W.T.tocsr()).tocsr()


def jaccard_connectivities(knn_indices, n, n_neighbors):
    """_neighbors/_neighbors.py:207-252: shared-neighbour Jaccard index of
    every directed edge (self excluded), symmetrised by averaging."""
    kk = n_neighbors - 1
    rows = np.repeat(np.arange(n), kk)
    cols = knn_indices[:, 1:].ravel()
    A = sparse.csr_matrix((np.ones(n * kk), (rows, cols)), shape=(n, n))
    shared = np.asarray(A[rows].multiply(A[cols]).sum(axis=1)).ravel()
    jac = shared / (2 * kk - shared)
    m = jac != 0
    W = sparse.coo_matrix((jac[m], (rows[m], cols[m])), shape=(n, n)).tocsr()
    return ((W + W.T) / 2).tocsr()
