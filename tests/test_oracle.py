"""Pins the CPU oracle (``oracle/harmony_oracle.py``) against checks that are
not restatements of its own formulas.

Each oracle function is checked by at least one of:

* the reference's own known-answer tests, with their hand-written inputs and
  expected values (``tests/test_harmony.py:342-450`` for lambda_kb) or their
  NumPy-seeded inputs (``tests/test_harmony.py:262-300`` scatter,
  ``tests/test_harmony_kernels.py:376-410`` inverse, which use
  ``np.random.default_rng(42)`` and are reproducible without CuPy);
* a textbook routine that computes the same quantity another way
  (``np.linalg.inv`` / ``np.linalg.solve`` of the explicit ridge system,
  ``scipy.special.softmax``, scikit-learn's Lloyd k-means, pandas
  categoricals, ``np.lexsort``);
* brute force on tiny inputs (pure-Python loops, different association);
* closed forms / invariants (theta=0 reduces the penalty to 1 and the row
  update to a softmax; ``sum_b O[b,k] = colsum(R)``; a bijective hash).

The k-means initialiser is pinned only as a textbook Lloyd iteration; parity
with the reference's cuML ``KMeans(init="k-means||")`` is unpinned (cuML is
not available), which is why every parity run injects ``Y0``.
"""

from __future__ import annotations

import math

import numpy as np
import pandas as pd
import pytest

from oracle import harmony_oracle as ho

DTYPES = [np.float32, np.float64]


# --------------------------------------------------------------------------
# lambda_kb: the reference's known-answer tests (tests/test_harmony.py:342-450)
# --------------------------------------------------------------------------


@pytest.mark.parametrize("dtype", DTYPES)
def test_lambda_kb_pruning_kat(dtype):
    # tests/test_harmony.py:342-383
    sentinel = dtype(ho.SUPPRESS_PENALTY)
    N_b = np.array([0, 100, 1, 50], dtype=dtype)
    O = np.array([[0, 0, 0], [30, 40, 30], [0, 0, 1], [20, 15, 15]], dtype=dtype)
    E = np.ones((4, 3), dtype=dtype) * 10
    r = ho.compute_lambda_kb(E, O, N_b, alpha=0.2, threshold=1e-5, ridge_lambda=1.0,
                             dynamic_lambda=True)
    assert r.dtype == dtype
    assert np.all(r[0] == sentinel)
    np.testing.assert_allclose(r[1], np.full(3, 0.2 * 10, dtype=dtype))
    assert r[2, 0] == sentinel and r[2, 1] == sentinel
    np.testing.assert_allclose(r[2, 2], dtype(0.2 * 10))
    np.testing.assert_allclose(r[3], np.full(3, 0.2 * 10, dtype=dtype))


@pytest.mark.parametrize("dtype", DTYPES)
def test_lambda_kb_dynamic_false_kat(dtype):
    # tests/test_harmony.py:386-403
    E = np.ones((3, 5), dtype=dtype)
    r = ho.compute_lambda_kb(E, np.ones((3, 5), dtype=dtype), np.ones(3, dtype=dtype),
                             alpha=0.5, threshold=1e-5, ridge_lambda=1.0,
                             dynamic_lambda=False)
    np.testing.assert_array_equal(r, np.full_like(E, 1.0))


@pytest.mark.parametrize("dtype", DTYPES)
def test_lambda_kb_fixed_ridge_zero_kat(dtype):
    # tests/test_harmony.py:406-426
    r = ho.compute_lambda_kb(np.array([[5.0, 0.0]], dtype=dtype),
                             np.array([[10.0, 0.0]], dtype=dtype),
                             np.array([100.0], dtype=dtype), alpha=0.2, threshold=None,
                             ridge_lambda=0.0, dynamic_lambda=False)
    assert r[0, 0] == dtype(0.0)
    assert r[0, 1] == dtype(ho.SUPPRESS_PENALTY)


@pytest.mark.parametrize("dtype", DTYPES)
def test_lambda_kb_zero_denominator_kat(dtype):
    # tests/test_harmony.py:429-450
    r = ho.compute_lambda_kb(np.array([[0.0, 5.0]], dtype=dtype),
                             np.array([[0.0, 10.0]], dtype=dtype),
                             np.array([100.0], dtype=dtype), alpha=0.2, threshold=None,
                             ridge_lambda=1.0, dynamic_lambda=True)
    assert r[0, 0] == dtype(ho.SUPPRESS_PENALTY)
    np.testing.assert_allclose(r[0, 1], dtype(1.0))


# --------------------------------------------------------------------------
# Scatter (O bookkeeping): exact integer test (tests/test_harmony.py:262-300)
# --------------------------------------------------------------------------


@pytest.mark.parametrize("n_cells", [1000, 60000])
@pytest.mark.parametrize("n_pcs", [20, 50])
@pytest.mark.parametrize("n_batches", [3, 10])
def test_scatter_rows_exact_integers(n_cells, n_pcs, n_batches):
    rng = np.random.default_rng(42)
    X = rng.integers(1, 10, size=(n_cells, n_pcs)).astype(np.float32)
    cats = rng.integers(0, n_batches, size=n_cells, dtype=np.int32)
    # independent: per-group integer sums (exact in int64, then exact in f32)
    expected = np.zeros((n_batches, n_pcs), dtype=np.int64)
    np.add.at(expected, cats, X.astype(np.int64))
    got = ho._scatter_rows(cats, X, n_batches)
    np.testing.assert_array_equal(got, expected.astype(np.float32))


# --------------------------------------------------------------------------
# Closed-form inverse vs np.linalg.inv (tests/test_harmony_kernels.py:376-410)
# --------------------------------------------------------------------------


def _explicit_system(O_col, lam_col):
    B = len(O_col)
    A = np.zeros((B + 1, B + 1), dtype=np.float64)
    A[0, 0] = O_col.sum()
    A[0, 1:] = O_col
    A[1:, 0] = O_col
    A[1:, 1:] = np.diag(O_col + lam_col)
    return A


@pytest.mark.parametrize("n_batches", [5, 50, 200])
@pytest.mark.parametrize("n_clusters", [10, 50])
@pytest.mark.parametrize("dtype", DTYPES)
def test_inv_mats_vs_linalg_inv(n_batches, n_clusters, dtype):
    rng = np.random.default_rng(42)
    O = (rng.random((n_batches, n_clusters)) * 100).astype(dtype)
    lam = (rng.random((n_batches, n_clusters)) * 5 + 0.1).astype(dtype)
    lam[0, 0] = dtype(1e30)  # pruned (batch, cluster) pair
    inv = ho.inv_mats(O, lam)
    assert inv.shape == (n_clusters, n_batches + 1, n_batches + 1)
    tol = 1e-4 if dtype == np.float32 else 1e-10
    for k in range(min(n_clusters, 3)):
        A = _explicit_system(O[:, k].astype(np.float64), lam[:, k].astype(np.float64))
        ref = np.linalg.inv(A)
        scale = np.abs(ref).max()
        np.testing.assert_allclose(inv[k], ref, atol=tol * scale, rtol=tol)
        # inverse property on the well-conditioned block
        eye = inv[k].astype(np.float64) @ A
        np.testing.assert_allclose(eye, np.eye(n_batches + 1), atol=1e3 * tol)


# --------------------------------------------------------------------------
# L2 / L1 normalisation (tests/test_harmony_kernels.py:54-86)
# --------------------------------------------------------------------------


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("shape", [(100, 50), (1, 20), (500, 3)])
def test_l2_normalize_vs_linalg_norm(dtype, shape):
    src = np.random.default_rng(42).standard_normal(shape).astype(dtype)
    expected = src / np.maximum(np.linalg.norm(src.astype(np.float64), axis=1,
                                               keepdims=True), 1e-12)
    atol = 1e-6 if dtype == np.float32 else 1e-12
    np.testing.assert_allclose(ho.l2_normalize(src), expected, atol=atol, rtol=1e-5)


@pytest.mark.parametrize("dtype", DTYPES)
def test_l2_normalize_zero_rows(dtype):
    src = np.zeros((3, 10), dtype=dtype)
    src[1] = 1.0
    dst = ho.l2_normalize(src)
    assert not np.any(np.isnan(dst))
    np.testing.assert_array_equal(dst[0], 0)
    np.testing.assert_array_equal(dst[2], 0)
    np.testing.assert_allclose(dst[1], 1 / math.sqrt(10), rtol=1e-6)


def test_l1_normalize_uses_absolute_values_and_clamp():
    X = np.array([[1.0, -3.0], [0.0, 0.0], [2.0, 2.0]])
    got = ho.l1_normalize(X)
    np.testing.assert_allclose(got[0], [0.25, -0.75])
    np.testing.assert_array_equal(got[1], [0.0, 0.0])
    np.testing.assert_allclose(got[2], [0.5, 0.5])


# --------------------------------------------------------------------------
# Batch bookkeeping (integer, exact)
# --------------------------------------------------------------------------


def test_category_index_mapping_vs_lexsort():
    rng = np.random.default_rng(7)
    for n, B in [(0, 3), (1, 1), (23, 4), (1024, 10)]:
        cats = rng.integers(0, B, size=n, dtype=np.int32)
        off, idx = ho.category_index_mapping(cats, B)
        lex = np.lexsort((np.arange(n), cats))
        np.testing.assert_array_equal(idx, lex)
        assert off[0] == 0 and off[-1] == n
        for b in range(B):
            seg = idx[off[b]:off[b + 1]]
            assert np.all(cats[seg] == b)
            assert np.all(np.diff(seg) > 0)


def test_category_index_mapping_hand_example():
    off, idx = ho.category_index_mapping(np.array([2, 0, 2, 1, 0], np.int32), 4)
    np.testing.assert_array_equal(off, [0, 2, 3, 5, 5])
    np.testing.assert_array_equal(idx, [1, 4, 3, 0, 2])


def test_batch_codes_single_and_combined_keys():
    obs = pd.DataFrame({"donor": ["b", "a", "c", "a"], "lane": [2, 1, 1, 2]})
    codes, cats, nb = ho.batch_codes(obs, "donor")
    assert cats == ["a", "b", "c"]
    np.testing.assert_array_equal(codes, [1, 0, 2, 0])
    np.testing.assert_array_equal(nb, [2, 1, 1])
    codes, cats, nb = ho.batch_codes(obs, ["donor", "lane"])
    assert cats == ["a,1", "a,2", "b,2", "c,1"]
    np.testing.assert_array_equal(codes, [2, 0, 3, 1])
    np.testing.assert_array_equal(nb, [1, 1, 1, 1])
    # an existing categorical keeps its unused categories (N_b = 0)
    obs2 = pd.DataFrame({"b": pd.Categorical(["x", "x"], categories=["x", "y"])})
    codes, cats, nb = ho.batch_codes(obs2, "b")
    assert cats == ["x", "y"]
    np.testing.assert_array_equal(nb, [2, 0])


def test_theta_tau_precedence_kat():
    # tau discount: theta * (1 - exp(-N_b/(K*tau))**2); pick N_b = K*tau*ln2 so
    # that exp(...) = 1/2 -> factor 3/4 (squaring the argument would give
    # 1 - exp(-ln2^2) = 0.3814...)
    K, tau = 10, 5.0
    nb = np.array([K * tau * math.log(2), 1e9])
    th = ho.theta_array(2.0, 2, np.float64, tau=tau, n_b=nb, n_clusters=K)
    np.testing.assert_allclose(th, [1.5, 2.0], rtol=1e-12)
    with pytest.raises(ValueError, match="must match"):
        ho.theta_array([1.0, 2.0, 3.0], 2, np.float64)
    np.testing.assert_array_equal(ho.theta_array(3, 4, np.float32), np.full(4, 3.0))


# --------------------------------------------------------------------------
# PCG keys and block order
# --------------------------------------------------------------------------


def _pcg_scalar(x: int) -> int:
    """The published PCG-RXS-M-XS 32-bit hash on Python integers (mod 2**32)."""
    m = 0xFFFFFFFF
    state = (x * 747796405 + 2891336453) & m
    word = (((state >> ((state >> 28) + 4)) ^ state) * 277803737) & m
    return ((word >> 22) ^ word) & m


@pytest.mark.parametrize("seed", [0, 1, 42, 1000003, 0xFFFFFFFF, 2**31 + 7])
def test_pcg_hash_vs_scalar_integers(seed):
    n = 5000
    got = ho.pcg_hash(n, seed)
    assert got.dtype == np.uint32
    want = np.array([_pcg_scalar(i ^ seed) for i in range(n)], dtype=np.uint64)
    np.testing.assert_array_equal(got.astype(np.uint64), want)


@pytest.mark.parametrize("seed", [0, 12345, 0xDEADBEEF])
def test_pcg_keys_are_distinct_and_order_is_permutation(seed):
    n = 1 << 20
    keys = ho.pcg_hash(n, seed)
    assert len(np.unique(keys)) == n  # the hash is a bijection on uint32
    order = ho.block_order(n, seed)
    assert np.array_equal(np.sort(order), np.arange(n))
    assert np.all(np.diff(keys[order].astype(np.int64)) > 0)


def test_clustering_seed_wraps_uint32():
    assert ho.clustering_seed(0, 0, 0) == 0
    assert ho.clustering_seed(0, 3, 5) == 3 * 1000003 + 5
    assert ho.clustering_seed(0xFFFFFFFF, 0, 1) == 0
    assert ho.clustering_seed(2**32 - 2, 1, 0) == (2**32 - 2 + 1000003) % 2**32


# --------------------------------------------------------------------------
# Penalty and objective
# --------------------------------------------------------------------------


@pytest.mark.parametrize("stabilized", [True, False])
def test_penalty_special_cases(stabilized):
    rng = np.random.default_rng(123)
    E = rng.random((3, 20)) * 10
    O = rng.random((3, 20)) * 10
    # theta = 0 -> no penalty
    np.testing.assert_array_equal(ho.penalty(E, O, np.zeros(3), stabilized), 1.0)
    # theta multiplicative: pen(t1 + t2) = pen(t1) * pen(t2)
    t1, t2 = rng.random(3), rng.random(3)
    np.testing.assert_allclose(ho.penalty(E, O, t1 + t2, stabilized),
                               ho.penalty(E, O, t1, stabilized) * ho.penalty(E, O, t2, stabilized),
                               rtol=1e-12)
    # one hand value per form: E=3, O=1, theta=2
    e, o = np.array([[3.0]]), np.array([[1.0]])
    want = (4 / 5) ** 2 if stabilized else (4 / 2) ** 2
    np.testing.assert_allclose(ho.penalty(e, o, np.array([2.0]), stabilized), want)


def _objective_loops(R, sim, O, E, theta, sigma, stabilized):
    n, K = R.shape
    km = 0.0
    ent = 0.0
    for i in range(n):
        rs = sum(R[i])
        for k in range(K):
            km += R[i, k] * 2.0 * (1.0 - sim[i, k])
            x = R[i, k] / rs
            ent += x * math.log(x + 1e-12)
    div = 0.0
    for b in range(O.shape[0]):
        for k in range(K):
            numer = O[b, k] + E[b, k] + 1 if stabilized else O[b, k] + 1
            div += theta[b] * O[b, k] * math.log(numer / (E[b, k] + 1))
    return km + sigma * ent + sigma * div


@pytest.mark.parametrize("stabilized", [True, False])
@pytest.mark.parametrize("n,K,B", [(50, 7, 3), (20, 5, 2)])
def test_objective_vs_loops(stabilized, n, K, B):
    rng = np.random.default_rng(42)
    R = rng.random((n, K)) + 0.01
    sim = rng.random((n, K))
    O = rng.random((B, K)) * 10 + 0.1
    E = rng.random((B, K)) * 10 + 0.1
    theta = rng.random(B) * 2
    got = ho.objective(R, sim, O, E, theta, 0.1, stabilized)
    want = _objective_loops(R, sim, O, E, theta, 0.1, stabilized)
    np.testing.assert_allclose(got, want, rtol=1e-12)


# --------------------------------------------------------------------------
# Clustering: invariants, theta=0 closed form, brute force on a tiny case
# --------------------------------------------------------------------------


def _small_problem(n=300, d=6, B=3, K=4, seed=0):
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((n, d)) + rng.standard_normal((K, d))[rng.integers(0, K, n)] * 2
    cats = rng.integers(0, B, n).astype(np.int32)
    Z = ho.l2_normalize(X)
    Y0 = Z[rng.choice(n, K, replace=False)]
    Nb = np.bincount(cats, minlength=B).astype(np.float64)
    Pr = (Nb / n).reshape(-1, 1)
    return X, cats, Z, Y0, Pr, Nb


def test_init_clusters_closed_form():
    from scipy.special import softmax

    X, cats, Z, Y0, Pr, _ = _small_problem()
    th = np.full(3, 2.0)
    R, E, O, obj = ho.init_clusters(Z, Y0, sigma=0.1, Pr_b=Pr, theta=th, cats=cats,
                                    n_batches=3, stabilized=True)
    Yn = Y0 / np.linalg.norm(Y0, axis=1, keepdims=True)
    np.testing.assert_allclose(R, softmax(-20.0 * (1 - Z @ Yn.T), axis=1), rtol=1e-10)
    np.testing.assert_allclose(O.sum(axis=0), R.sum(axis=0), rtol=1e-12)
    np.testing.assert_allclose(E, Pr * R.sum(axis=0)[None, :], rtol=1e-12)
    for b in range(3):
        np.testing.assert_allclose(O[b], R[cats == b].sum(axis=0), rtol=1e-12)


def test_clustering_theta_zero_is_softmax():
    """theta = 0 -> penalty 1 -> every row becomes the softmax of its
    start-of-iteration similarities, whatever the block order."""
    from scipy.special import softmax

    X, cats, Z, Y0, Pr, _ = _small_problem()
    th = np.zeros(3)
    R, E, O, _ = ho.init_clusters(Z, Y0, sigma=0.1, Pr_b=Pr, theta=th, cats=cats,
                                  n_batches=3, stabilized=False)
    Y = R.T @ Z
    Yn = Y / np.linalg.norm(Y, axis=1, keepdims=True)
    want = softmax(-20.0 * (1 - Z @ Yn.T), axis=1)
    ho.clustering_loop(Z, R, E, O, Pr, cats, th, sigma=0.1, tol=-np.inf, max_iter=1,
                       block_size=17, seed=5, stabilized=False)
    np.testing.assert_allclose(R, want, rtol=1e-10)


@pytest.mark.parametrize("stabilized", [True, False])
def test_clustering_invariants(stabilized):
    X, cats, Z, Y0, Pr, _ = _small_problem()
    th = np.full(3, 2.0)
    R, E, O, obj0 = ho.init_clusters(Z, Y0, sigma=0.1, Pr_b=Pr, theta=th, cats=cats,
                                     n_batches=3, stabilized=stabilized)
    last, nit, objs = ho.clustering_loop(Z, R, E, O, Pr, cats, th, sigma=0.1, tol=-np.inf,
                                         max_iter=4, block_size=15, seed=9,
                                         stabilized=stabilized)
    assert nit == 4 and len(objs) == 4 and last == objs[-1]
    np.testing.assert_allclose(R.sum(axis=1), 1.0, rtol=1e-12)
    for b in range(3):
        np.testing.assert_allclose(O[b], R[cats == b].sum(axis=0), rtol=1e-10, atol=1e-12)
    np.testing.assert_allclose(E, Pr * R.sum(axis=0)[None, :], rtol=1e-10, atol=1e-12)


def _clustering_brute(Z, R, cats, Pr, theta, sigma, block, seed, stabilized):
    """One clustering iteration with explicit per-cell loops."""
    n, K = R.shape
    B = len(Pr)
    Y = [[sum(R[i, k] * Z[i, j] for i in range(n)) for j in range(Z.shape[1])]
         for k in range(K)]
    Yn = [[v / math.sqrt(sum(u * u for u in row)) for v in row] for row in Y]
    sim = [[sum(Z[i, j] * Yn[k][j] for j in range(Z.shape[1])) for k in range(K)]
           for i in range(n)]
    O = [[sum(R[i, k] for i in range(n) if cats[i] == b) for k in range(K)] for b in range(B)]
    E = [[Pr[b] * sum(R[i, k] for i in range(n)) for k in range(K)] for b in range(B)]
    keys = [_pcg_scalar(i ^ seed) for i in range(n)]
    order = sorted(range(n), key=lambda i: keys[i])
    R = R.copy()
    for s in range(0, n, block):
        idx = order[s:s + block]
        for i in idx:
            for k in range(K):
                O[cats[i]][k] -= R[i, k]
                for b in range(B):
                    E[b][k] -= Pr[b] * R[i, k]
        pen = [[((E[b][k] + 1) / ((O[b][k] + E[b][k] + 1) if stabilized else (O[b][k] + 1)))
                ** theta[b] for k in range(K)] for b in range(B)]
        for i in idx:
            v = [math.exp(-2 / sigma * (1 - sim[i][k])) * pen[cats[i]][k] for k in range(K)]
            tot = sum(v)
            for k in range(K):
                R[i, k] = v[k] / tot
                O[cats[i]][k] += R[i, k]
                for b in range(B):
                    E[b][k] += Pr[b] * R[i, k]
    return R, np.array(O), np.array(E)


@pytest.mark.parametrize("stabilized", [True, False])
def test_clustering_brute_force_tiny(stabilized):
    rng = np.random.default_rng(3)
    n, d, K, B = 9, 3, 2, 2
    Z = ho.l2_normalize(rng.standard_normal((n, d)))
    cats = np.array([0, 1, 0, 0, 1, 1, 0, 1, 0], np.int32)
    Pr = np.bincount(cats, minlength=B) / n
    th = np.array([2.0, 0.7])
    R0 = ho.l1_normalize(rng.random((n, K)) + 0.1)
    R = R0.copy()
    O = ho._scatter_rows(cats, R, B)
    E = Pr[:, None] * R.sum(axis=0)[None, :]
    ho.clustering_loop(Z, R, E, O, Pr.reshape(-1, 1), cats, th, sigma=0.3, tol=-np.inf,
                       max_iter=1, block_size=2, seed=11, stabilized=stabilized)
    Rb, Ob, Eb = _clustering_brute(Z, R0, cats, Pr, th, 0.3, 2, 11, stabilized)
    np.testing.assert_allclose(R, Rb, rtol=1e-12)
    np.testing.assert_allclose(O, Ob, rtol=1e-12)
    np.testing.assert_allclose(E, Eb, rtol=1e-12)


def test_clustering_window3_rule():
    """The loop stops at the first iteration (>= 4 objectives) whose window of
    the last three objectives fails to improve by tol*|old window|."""
    X, cats, Z, Y0, Pr, _ = _small_problem(n=200, seed=4)
    th = np.full(3, 2.0)
    R, E, O, _ = ho.init_clusters(Z, Y0, sigma=0.1, Pr_b=Pr, theta=th, cats=cats,
                                  n_batches=3, stabilized=True)
    R2, E2, O2 = R.copy(), E.copy(), O.copy()
    _, n_full, objs = ho.clustering_loop(Z, R2, E2, O2, Pr, cats, th, sigma=0.1,
                                         tol=-np.inf, max_iter=12, block_size=10,
                                         seed=1, stabilized=True)
    tol = 1e-3
    stop = 12
    for i in range(3, 12):
        old = sum(objs[i - 3:i])
        new = sum(objs[i - 2:i + 1])
        if old - new < tol * abs(old):
            stop = i + 1
            break
    _, n_it, _ = ho.clustering_loop(Z, R, E, O, Pr, cats, th, sigma=0.1, tol=tol,
                                    max_iter=12, block_size=10, seed=1, stabilized=True)
    assert n_it == stop


# --------------------------------------------------------------------------
# Correction: closed form vs explicit inverse vs dense ridge solve
# --------------------------------------------------------------------------


def _ridge_dense(X, R, lam, cats, B):
    """Textbook ridge regression per cluster on the full design matrix."""
    n = X.shape[0]
    D = np.zeros((n, B + 1))
    D[:, 0] = 1
    D[np.arange(n), cats + 1] = 1
    Z = X.copy()
    for k in range(R.shape[1]):
        w = R[:, k]
        A = D.T @ (w[:, None] * D) + np.diag(np.r_[0.0, lam[:, k]])
        Wk = np.linalg.solve(A, D.T @ (w[:, None] * X))
        Wk[0] = 0
        Z -= w[:, None] * (D @ Wk)
    return Z


@pytest.mark.parametrize("dynamic", [False, True])
def test_correction_methods_vs_dense_solve(dynamic):
    X, cats, Z, Y0, Pr, Nb = _small_problem(n=400, d=5, B=3, K=4, seed=2)
    th = np.full(3, 2.0)
    R, E, O, _ = ho.init_clusters(Z, Y0, sigma=0.1, Pr_b=Pr, theta=th, cats=cats,
                                  n_batches=3, stabilized=True)
    lam = ho.compute_lambda_kb(E, O, Nb, alpha=0.2, threshold=1e-5, ridge_lambda=1.0,
                               dynamic_lambda=dynamic)
    want = _ridge_dense(X, R, lam, cats, 3)
    zb, W = ho.correction_batched(X, R, O, lam, cats, 3)
    zo = ho.correction_original(X, R, lam, cats, 3)
    np.testing.assert_allclose(zb, want, rtol=1e-9, atol=1e-10)
    np.testing.assert_allclose(zo, want, rtol=1e-9, atol=1e-10)
    assert np.all(W[:, 0, :] == 0)


def test_correction_tiny_brute_force():
    """N=6, K=2, B=2 with explicit np.linalg.solve per cluster."""
    X = np.array([[1.0, 2.0], [0.5, -1.0], [3.0, 0.0], [-1.0, 1.0], [2.0, 2.0], [0.0, 1.0]])
    cats = np.array([0, 1, 0, 1, 1, 0], np.int32)
    R = np.array([[0.9, 0.1], [0.2, 0.8], [0.5, 0.5], [0.3, 0.7], [0.6, 0.4], [1.0, 0.0]])
    O = ho._scatter_rows(cats, R, 2)
    lam = np.full((2, 2), 1.0)
    want = _ridge_dense(X, R, lam, cats, 2)
    zb, _ = ho.correction_batched(X, R, O, lam, cats, 2)
    np.testing.assert_allclose(zb, want, rtol=1e-12, atol=1e-13)


def test_correction_suppressed_pair_gets_no_batch_term():
    """lambda = 1e30 for (b,k) -> W_k[b+1] ~ 0, so cells of batch b receive no
    batch-b correction from cluster k."""
    X, cats, Z, Y0, Pr, Nb = _small_problem(n=200, d=4, B=2, K=2, seed=8)
    R, E, O, _ = ho.init_clusters(Z, Y0, sigma=0.1, Pr_b=Pr, theta=np.full(2, 2.0),
                                  cats=cats, n_batches=2, stabilized=True)
    lam = np.ones((2, 2))
    lam[1, 0] = ho.SUPPRESS_PENALTY
    _, W = ho.correction_batched(X, R, O, lam, cats, 2)
    assert np.abs(W[0, 2]).max() < 1e-20
    assert np.abs(W[0, 1]).max() > 1e-3


# --------------------------------------------------------------------------
# K-means initialiser: textbook Lloyd (scikit-learn) from the same seeds
# --------------------------------------------------------------------------


def test_kmeans_is_lloyd():
    from sklearn.cluster import KMeans

    rng = np.random.default_rng(0)
    centers = rng.standard_normal((6, 8)) * 4
    X = centers[rng.integers(0, 6, 3000)] + rng.standard_normal((3000, 8))
    Z = ho.l2_normalize(X)
    Y = ho.kmeans_centroids(Z, 6, random_state=3, max_iter=25)
    seeds = ho.block_order(3000, 3)[:6]
    km = KMeans(n_clusters=6, init=Z[seeds], n_init=1, max_iter=25, tol=0.0,
                algorithm="lloyd").fit(Z)
    np.testing.assert_allclose(Y, km.cluster_centers_, rtol=1e-9, atol=1e-12)


# --------------------------------------------------------------------------
# Whole run: Harmony convergence rule and a golden regression vector
# --------------------------------------------------------------------------


def test_harmonize_convergence_uses_last_two_objectives():
    X, cats, Z, Y0, Pr, _ = _small_problem(n=300, seed=5)
    rec = {}
    ho.harmonize(X, cats, 3, Y0=Y0, n_clusters=4, max_iter_harmony=10,
                 max_iter_clustering=5, tol_harmony=1e-2, record=rec)
    objs = rec["objectives"]
    n_h = len(objs) - 1
    for i in range(1, n_h):  # no earlier stop was due
        assert objs[i - 1] - objs[i] >= 1e-2 * abs(objs[i - 1])
    if n_h < 10:
        assert objs[-2] - objs[-1] < 1e-2 * abs(objs[-2])


# --------------------------------------------------------------------------
# Several covariates (stacked one-hot Phi, build extension).  Pinned by two
# reductions to the one-covariate oracle (itself pinned above) and by the
# invariants of stacked one-hot bookkeeping.
# --------------------------------------------------------------------------


def _multi_problem(n=600, d=6, B=3, D=4, seed=11):
    rng = np.random.default_rng(seed)
    centers = rng.standard_normal((5, d)) * 3
    b = rng.integers(0, B, n)
    dn = rng.integers(0, D, n)
    X = (centers[rng.integers(0, 5, n)] + rng.standard_normal((n, d))
         + rng.standard_normal((B, d))[b] + 0.5 * rng.standard_normal((D, d))[dn])
    Z = ho.l2_normalize(X)
    Y0 = Z[rng.choice(n, 6, replace=False)]
    return X, b.astype(np.int32), dn.astype(np.int32), Y0


MULTI_KW = dict(n_clusters=6, max_iter_harmony=3, max_iter_clustering=6, tol_harmony=-np.inf,
                tol_clustering=-np.inf, random_state=5, dynamic_lambda=False)


@pytest.mark.parametrize("stabilized", [False, True])
def test_multi_one_covariate_is_the_reference_algorithm(stabilized):
    X, b, _, Y0 = _multi_problem()
    r1, r2 = {}, {}
    Z1 = ho.harmonize(X, b, 3, Y0=Y0, stabilized_penalty=stabilized, record=r1, **MULTI_KW)
    Z2 = ho.harmonize_multi(X, [b], [3], Y0=Y0, stabilized_penalty=stabilized, record=r2,
                            **MULTI_KW)
    np.testing.assert_allclose(r2["objectives"], r1["objectives"], rtol=1e-12)
    np.testing.assert_allclose(r2["R"], r1["R"], rtol=1e-9, atol=1e-13)
    assert np.linalg.norm(Z2 - Z1) / np.linalg.norm(Z1) < 1e-10


@pytest.mark.parametrize("stabilized", [False, True])
def test_multi_duplicated_covariate_doubles_theta_and_halves_lambda(stabilized):
    """A covariate given twice: its two levels per cell multiply the same
    penalty (theta -> 2 theta) and the ridge solution splits the batch term
    evenly over the two copies (lambda -> lambda / 2)."""
    X, b, _, Y0 = _multi_problem(seed=3)
    r1, r2 = {}, {}
    kw = dict(MULTI_KW)
    Z1 = ho.harmonize(X, b, 3, Y0=Y0, theta=3.0, ridge_lambda=0.5, stabilized_penalty=stabilized,
                      record=r1, **kw)
    Z2 = ho.harmonize_multi(X, [b, b], [3, 3], Y0=Y0, theta=1.5, ridge_lambda=1.0,
                            stabilized_penalty=stabilized, record=r2, **kw)
    np.testing.assert_allclose(r2["objectives"], r1["objectives"], rtol=1e-10)
    np.testing.assert_allclose(r2["R"], r1["R"], rtol=1e-8, atol=1e-12)
    assert np.linalg.norm(Z2 - Z1) / np.linalg.norm(Z1) < 1e-9


def test_multi_bookkeeping_invariants():
    X, b, dn, Y0 = _multi_problem(seed=9)
    rec = {}
    ho.harmonize_multi(X, [b, dn], [3, 4], Y0=Y0, record=rec, **MULTI_KW)
    R, O, E = rec["R"], rec["O"], rec["E"]
    colsum = R.sum(axis=0)
    # every covariate's levels partition the cells
    np.testing.assert_allclose(O[:3].sum(axis=0), colsum, rtol=1e-10)
    np.testing.assert_allclose(O[3:].sum(axis=0), colsum, rtol=1e-10)
    np.testing.assert_allclose(O[:3], ho._scatter_rows(b, R, 3), rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(O[3:], ho._scatter_rows(dn, R, 4), rtol=1e-9, atol=1e-12)
    nb = np.concatenate([np.bincount(b, minlength=3), np.bincount(dn, minlength=4)])
    np.testing.assert_allclose(E, (nb / len(b))[:, None] * colsum[None, :], rtol=1e-9)
    np.testing.assert_allclose(R.sum(axis=1), 1.0, rtol=1e-12)


def test_multi_penalty_is_product_over_levels_brute_force():
    """One block update on a 9-cell problem, penalties multiplied cell by cell
    in plain Python."""
    rng = np.random.default_rng(2)
    n, K = 9, 3
    lev = np.array([[rng.integers(0, 2), 2 + rng.integers(0, 3)] for _ in range(n)], np.int32)
    pen = rng.uniform(0.2, 2.0, size=(5, K))
    got = ho.multi_penalty_rows(pen, lev)
    for i in range(n):
        for k in range(K):
            assert got[i, k] == pytest.approx(pen[lev[i, 0], k] * pen[lev[i, 1], k], rel=1e-15)


def test_multi_correction_solves_the_normal_equations():
    X, b, dn, _ = _multi_problem(n=300, seed=4)
    rng = np.random.default_rng(0)
    R = ho.l1_normalize(rng.uniform(0.01, 1.0, size=(300, 4)))
    lev, _ = ho.multi_levels([b, dn], [3, 4])
    lam = np.full((7, 4), 0.7)
    Z, W = ho.multi_correction(X, R, lam, lev, 7)
    P = np.zeros((300, 8))
    P[:, 0] = 1
    P[np.arange(300), 1 + lev[:, 0]] = 1
    P[np.arange(300), 1 + lev[:, 1]] = 1
    for k in range(4):
        A = (P * R[:, k:k + 1]).T @ P + np.diag(np.concatenate([[0.0], lam[:, k]]))
        rhs = (P * R[:, k:k + 1]).T @ X
        full = np.linalg.lstsq(A, rhs, rcond=None)[0]
        np.testing.assert_allclose(W[k, 1:], full[1:], rtol=1e-8, atol=1e-10)
    expect = X - sum(R[:, k:k + 1] * (P[:, 1:] @ W[k, 1:]) for k in range(4))
    np.testing.assert_allclose(Z, expect, rtol=1e-10, atol=1e-12)
