"""The neighbour-graph oracle pinned against scikit-learn, the defining
equation of the UMAP normalisation and brute-force loops (CPU)."""

from __future__ import annotations

import numpy as np
import pytest
from scipy import sparse

from oracle import neighbors_oracle as no


def _data(n=400, d=12, seed=0):
    rng = np.random.default_rng(seed)
    centers = rng.normal(0, 3, size=(6, d))
    return centers[rng.integers(0, 6, n)] + rng.normal(size=(n, d))


@pytest.mark.parametrize("metric", ["euclidean", "sqeuclidean", "cosine"])
def test_knn_matches_sklearn(metric):
    from sklearn.neighbors import NearestNeighbors

    X = _data()
    idx, dist = no.knn_brute(X, 10, metric)
    skm = "euclidean" if metric == "sqeuclidean" else metric
    ref_d, ref_i = NearestNeighbors(n_neighbors=10, algorithm="brute", metric=skm).fit(X).kneighbors(X)
    if metric == "sqeuclidean":
        ref_d = ref_d ** 2
    # the self distance of the expanded form is a rounding residue (the
    # reference zeroes column 0 afterwards)
    assert np.abs(dist[:, 0]).max() < 1e-6
    np.testing.assert_allclose(dist[:, 1:], ref_d[:, 1:], rtol=1e-9, atol=1e-9)
    # same neighbour sets (continuous data: no ties)
    assert (np.sort(idx, axis=1) == np.sort(ref_i, axis=1)).all()
    assert (idx[:, 0] == np.arange(len(X))).all()


def test_knn_ties_broken_by_index():
    X = np.repeat(np.arange(5.0)[:, None], 3, axis=0)  # each point three times
    idx, dist = no.knn_brute(X, 3, "euclidean")
    assert (dist == 0).all()
    for i in range(len(X)):
        b = (i // 3) * 3
        assert list(idx[i]) == [b, b + 1, b + 2]


def test_smooth_knn_dist_solves_its_equation():
    X = _data(n=300)
    k = 15
    _, dist = no.knn_brute(X, k, "euclidean")
    sig, rho = no.smooth_knn_dist(dist, float(k))
    target = np.log2(k)
    floored = 0
    for i in range(len(X)):
        # rho is the nearest non-zero distance
        assert rho[i] == dist[i][dist[i] > 0].min()
        psum = np.sum(np.exp(-np.maximum(dist[i, 1:] - rho[i], 0) / sig[i]))
        if sig[i] <= 1e-3 * dist[i].mean() * (1 + 1e-12):
            floored += 1
            continue
        assert abs(psum - target) < 1e-4
    assert floored < len(X)


def test_membership_and_union_brute_force():
    rng = np.random.default_rng(3)
    n, k = 7, 3
    idx = np.stack([np.concatenate([[i], rng.choice([j for j in range(n) if j != i], k - 1,
                                                    replace=False)]) for i in range(n)])
    dist = np.sort(rng.uniform(0.1, 2.0, size=(n, k)), axis=1)
    dist[:, 0] = 0.0
    sig, rho = no.smooth_knn_dist(dist, float(k))
    vals = no.membership_strengths(idx, dist, sig, rho)
    G = no.fuzzy_union(idx, vals, n).toarray()
    W = np.zeros((n, n))
    for i in range(n):
        for j in range(k):
            c = idx[i, j]
            if c == i:
                v = 0.0
            elif dist[i, j] - rho[i] <= 0:
                v = 1.0
            else:
                v = np.exp(-(dist[i, j] - rho[i]) / sig[i])
            W[i, c] = v
    expect = W + W.T - W * W.T
    np.testing.assert_allclose(G, expect, rtol=1e-12, atol=1e-15)
    assert (np.diag(G) == 0).all() and np.allclose(G, G.T)
    assert G.max() <= 1.0 + 1e-12


def test_gauss_and_jaccard_brute_force():
    X = _data(n=60, d=4, seed=5)
    k = 6
    idx, dist = no.knn_brute(X, k, "euclidean")
    dist = no.fix_self_distances(dist, "euclidean")
    n = len(X)
    G = no.gauss_connectivities(idx, dist, n).toarray()
    sig = np.sqrt(np.median(dist[:, 1:] ** 2, axis=1))
    W = np.zeros((n, n))
    for i in range(n):
        for j in range(1, k):
            c = idx[i, j]
            den = sig[i] ** 2 + sig[c] ** 2
            W[i, c] = np.sqrt(2 * sig[i] * sig[c] / den) * np.exp(-dist[i, j] ** 2 / den)
    np.testing.assert_allclose(G, np.maximum(W, W.T), rtol=1e-12)
    J = no.jaccard_connectivities(idx, n, k).toarray()
    nb = [set(idx[i, 1:]) for i in range(n)]
    Wj = np.zeros((n, n))
    for i in range(n):
        for c in idx[i, 1:]:
            s = len(nb[i] & nb[c])
            Wj[i, c] = s / (2 * (k - 1) - s)
    np.testing.assert_allclose(J, (Wj + Wj.T) / 2, rtol=1e-12)


def test_sparse_distances_layout():
    X = _data(n=50)
    idx, dist = no.knn_brute(X, 5, "euclidean")
    D = no.sparse_distances(idx, dist, 50)
    assert isinstance(D, sparse.csr_matrix) and D.nnz == 250
    np.testing.assert_array_equal(D.indptr, np.arange(0, 251, 5))
