"""CPU oracle for the Harmony hot path -- TEST INFRASTRUCTURE ONLY.

This module is a NumPy restatement of rapids_singlecell's Harmony algorithm
(``rsc.pp.harmony_integrate``).  It is the checker against which the HIP
implementation in ``rapids_singlecell_amd`` is compared.  Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import it; the product path never does.

Every function names the reference file:line it restates (paths relative to
``src/rapids_singlecell/``).  Arithmetic is carried out in the dtype of the
inputs (float64 by default), mirroring the reference's templated kernels.

How each part is pinned (checked against something other than itself) is
listed in ``tests/test_oracle.py``; the one unpinned part is the k-means
initialiser (the reference delegates to cuML, which is absent here), which is
why every parity run injects the initial centroids ``Y0``.
"""

from __future__ import annotations

import numpy as np

# _harmony/__init__.py:38
SUPPRESS_PENALTY = 1e30
_U32 = 0xFFFFFFFF


def _scatter_rows(cats: np.ndarray, rows: np.ndarray, n_groups: int) -> np.ndarray:
    """sum of ``rows`` per group id (``np.add.at`` semantics, as one product)."""
    onehot = np.zeros((n_groups, len(cats)), dtype=rows.dtype)
    onehot[cats, np.arange(len(cats))] = 1
    return onehot @ rows


# --------------------------------------------------------------------------
# Row normalisation
# --------------------------------------------------------------------------


def l2_normalize(X: np.ndarray) -> np.ndarray:
    """L2 row normalise: ``x * min(rsqrt(sum x^2), 1e12)``.

    _cuda/harmony/normalize/kernels_normalize.cuh:6-54 (zero rows stay zero
    because rsqrt(0)=inf is clamped to 1e12).
    """
    X = np.ascontiguousarray(X)
    ss = np.einsum("ij,ij->i", X, X)
    with np.errstate(divide="ignore"):
        inv = np.where(ss > 0, 1.0 / np.sqrt(ss.astype(np.float64)), np.inf)
    inv = np.minimum(inv, 1e12).astype(X.dtype)
    return X * inv[:, None]


def l1_normalize(X: np.ndarray) -> np.ndarray:
    """L1 row normalise: ``x / max(sum |x|, 1e-12)``.

    _cuda/harmony/normalize/kernels_normalize.cuh:58-108.
    """
    s = np.abs(X).sum(axis=1)
    s = np.maximum(s, X.dtype.type(1e-12))
    return X * (X.dtype.type(1) / s)[:, None]


# --------------------------------------------------------------------------
# Batch bookkeeping (integer, bit-exact)
# --------------------------------------------------------------------------


def batch_codes(obs, key):
    """Categorical batch codes.

    preprocessing/_harmony/_helper.py:223-236: a single key is used as-is, a
    list of keys is joined row-wise with "," after ``astype(str)``; the result
    is ``astype("category")`` (pandas sorts the categories, unused categories
    of an existing categorical are kept).  Returns (codes int32, categories,
    counts N_b int64 in category order).
    """
    import pandas as pd  # noqa: F401  (pandas semantics are what we restate)

    if isinstance(key, str):
        vec = obs[key]
    elif len(key) == 1:
        vec = obs[key[0]]
    else:
        df = obs[list(key)].astype("str")
        vec = df.apply(lambda row: ",".join(row), axis=1)
    cat = vec.astype("category")
    codes = np.asarray(cat.cat.codes.values, dtype=np.int32)
    n_b = np.asarray(cat.value_counts(sort=False).values, dtype=np.int64)
    return codes, list(cat.cat.categories), n_b


def category_index_mapping(cats: np.ndarray, n_batches: int):
    """CSR of cells grouped by batch.

    _utils/_multi_gpu.py:202-232: ``cat_offsets`` is the exclusive cumsum of
    per-batch counts (length B+1); ``cell_indices`` is the stable argsort of
    ``cats`` (``lexsort((idx, cats))``).
    """
    counts = np.bincount(cats, minlength=n_batches).astype(np.int32)
    offsets = np.zeros(n_batches + 1, dtype=np.int32)
    np.cumsum(counts, out=offsets[1:])
    cell_indices = np.argsort(cats, kind="stable").astype(np.int32)
    return offsets, cell_indices


def theta_array(theta, n_batches: int, dtype, *, tau: float = 0, n_b=None,
                n_clusters: int | None = None) -> np.ndarray:
    """Per-batch theta (+ tau discount).

    preprocessing/_harmony/_helper.py:277-307 and _harmony/__init__.py:172-175.
    The tau discount is ``theta * (1 - exp(-N_b/(K*tau))**2)`` -- Python
    precedence squares the exponential, not its argument.
    """
    dtype = np.dtype(dtype)
    if isinstance(theta, (float, int)):  # bool is an int: True -> 1.0
        th = np.ones(n_batches, dtype=dtype) * float(theta)
    elif isinstance(theta, (list, np.ndarray)):
        th = np.asarray(theta, dtype=dtype)
        if th.size != n_batches:
            raise ValueError(
                f"Theta array size ({th.size}) must match number of batches "
                f"({n_batches})"
            )
        th = th.ravel()
    else:
        raise ValueError(
            "Theta must be float, int, list, numpy array, or cupy array, got "
            f"{type(theta)}"
        )
    if tau > 0:
        nb = np.asarray(n_b, dtype=dtype)
        th = th * (1 - np.exp(-nb / (n_clusters * tau)) ** 2)
    return np.ascontiguousarray(th.astype(dtype).ravel())


# --------------------------------------------------------------------------
# Block order: PCG hash + sort (integer, bit-exact)
# --------------------------------------------------------------------------


def pcg_hash(n: int, seed: int) -> np.ndarray:
    """Shuffle keys ``key_i = pcg(i ^ seed)`` as uint32.

    _cuda/harmony/clustering/kernels_clustering.cuh:106-117.
    """
    i = np.arange(n, dtype=np.uint32)
    state = (i ^ np.uint32(seed & _U32)) * np.uint32(747796405) + np.uint32(
        2891336453
    )
    word = ((state >> ((state >> np.uint32(28)) + np.uint32(4))) ^ state) * np.uint32(
        277803737
    )
    return (word >> np.uint32(22)) ^ word


def block_order(n: int, seed: int) -> np.ndarray:
    """Ascending sort of (key, i) -> permutation (``idx_list_alt``).

    _cuda/harmony/clustering/clustering.cu:292-305 (CUB radix sort of the
    pcg keys with iota values; keys are distinct, so the order is unique).
    """
    keys = pcg_hash(n, seed)
    return np.argsort(keys, kind="stable").astype(np.int32)


def clustering_seed(random_state: int, harmony_iter: int, clustering_iter: int) -> int:
    """Seed of one clustering iteration.

    _harmony/__init__.py:249 (``random_state + i*1000003``), :446
    (``& 0xFFFFFFFF``) and clustering.cu:293 (``seed + iter`` in uint32).
    """
    base = (random_state + harmony_iter * 1000003) & _U32
    return (base + clustering_iter) & _U32


# --------------------------------------------------------------------------
# Penalty, objective
# --------------------------------------------------------------------------


def penalty(E, O, theta, stabilized: bool):
    """``((E+1)/(O+E+1))^theta`` (harmony2) or ``((E+1)/(O+1))^theta``.

    _cuda/harmony/pen/kernels_pen.cuh:8-20.
    """
    one = E.dtype.type(1)
    denom = (O + E + one) if stabilized else (O + one)
    return np.power((E + one) / denom, theta[:, None])


def objective(R, sim, O, E, theta, sigma, stabilized: bool):
    """Three-term Harmony objective.

    clustering.cu:188-237; kmeans term kernels_kmeans.cuh:30-95 (``R*2*(1-sim)``);
    entropy kernels_clustering.cuh:9-62 (row-normalised R, eps 1e-12, times
    sigma); diversity kernels_clustering.cuh:69-103.
    """
    dt = R.dtype.type
    kmeans_err = np.sum(R * dt(2) * (dt(1) - sim), dtype=R.dtype)
    Rn = R / R.sum(axis=1, keepdims=True)
    entropy = dt(sigma) * np.sum(Rn * np.log(Rn + dt(1e-12)), dtype=R.dtype)
    numer = (O + E + dt(1)) if stabilized else (O + dt(1))
    diversity = dt(sigma) * np.sum(
        theta[:, None] * O * np.log(numer / (E + dt(1))), dtype=R.dtype
    )
    return dt(kmeans_err + entropy + diversity)


# --------------------------------------------------------------------------
# Initialisation (given centroids)
# --------------------------------------------------------------------------


def init_clusters(Z_norm, Y0, *, sigma, Pr_b, theta, cats, n_batches,
                  stabilized: bool):
    """R, E, O and the first objective from initial centroids ``Y0``.

    _harmony/__init__.py:323-353: ``Y_norm = L2(Y0)``; ``R = L1(exp(-2/sigma
    (1 - Z_norm Y_norm^T)))``; ``E = Pr_b (x) colsum(R)``; ``O = scatter(R by
    cat)``; objective from the init similarities.
    """
    dt = Z_norm.dtype
    Y_norm = l2_normalize(np.asarray(Y0, dtype=dt))
    term = dt.type(-2 / sigma)
    sim = Z_norm @ Y_norm.T
    R = np.exp(term * (dt.type(1) - sim))
    R = l1_normalize(R)
    E = Pr_b.reshape(-1, 1) * R.sum(axis=0)[None, :]
    O = _scatter_rows(cats, R, n_batches)
    obj = objective(R, sim, O, E, theta, sigma, stabilized)
    return R, E.astype(dt), O, float(obj)


# --------------------------------------------------------------------------
# Clustering loop (the hot loop)
# --------------------------------------------------------------------------


def clustering_loop(Z_norm, R, E, O, Pr_b, cats, theta, *, sigma, tol,
                    max_iter, block_size, seed, stabilized, trace=None):
    """One Harmony iteration's clustering step; mutates R, E, O in place.

    clustering.cu:241-393.  Per clustering iteration: centroids
    ``Y = R^T Z_norm`` (:272-276), ``Y_norm = L2(Y)`` (:278-281),
    ``sim = Z_norm Y_norm^T`` (:284-289), block order from ``seed + iter``
    (:292-305); then for each block of ``block_size`` rows in that order
    (:308-372): remove the block's old contribution from O and E, compute the
    penalty, ``R_out = L1(exp(term (1-sim)) * pen[cat])``, write back and add
    the new contribution.  The objective (:375-377) uses the final R and the
    start-of-iteration ``sim``; the window-3 test (:380-387) compares the sum
    of the last three objectives with the sum of the three before.

    Returns ``(last_obj, n_iter, objectives)``.
    """
    dt = R.dtype.type
    n, _ = R.shape
    term = dt(-2) / dt(sigma)
    tol_t = dt(tol)
    objectives = []
    Pr = Pr_b.reshape(-1)
    for it in range(max_iter):
        Y = R.T @ Z_norm
        Y_norm = l2_normalize(Y)
        sim = Z_norm @ Y_norm.T
        order = block_order(n, (seed + it) & _U32)
        for pos in range(0, n, block_size):
            idx = order[pos : pos + block_size]
            R_in = R[idx]
            c_in = cats[idx]
            R_in_sum = R_in.sum(axis=0)
            O -= _scatter_rows(c_in, R_in, O.shape[0])
            E -= Pr[:, None] * R_in_sum[None, :]
            pen = penalty(E, O, theta, stabilized)
            v = np.exp(term * (dt(1) - sim[idx])) * pen[c_in]
            R_out = v * (dt(1) / v.sum(axis=1))[:, None]
            R[idx] = R_out
            R_out_sum = R_out.sum(axis=0)
            O += _scatter_rows(c_in, R_out, O.shape[0])
            E += Pr[:, None] * R_out_sum[None, :]
        obj = objective(R, sim, O, E, theta, sigma, stabilized)
        objectives.append(obj)
        if trace is not None:
            trace.append({"iter": it, "Y_norm": Y_norm.copy(), "obj": obj})
        if len(objectives) >= 4:
            old = dt(0)
            new = dt(0)
            for w in range(3):
                old = dt(old + objectives[-2 - w])
                new = dt(new + objectives[-1 - w])
            if (old - new) < tol_t * abs(old):
                break
    last = objectives[-1] if objectives else dt(0)
    return float(last), len(objectives), [float(o) for o in objectives]


# --------------------------------------------------------------------------
# Ridge correction
# --------------------------------------------------------------------------


def compute_lambda_kb(E, O, N_b, *, alpha, threshold, ridge_lambda,
                      dynamic_lambda: bool):
    """Per-(batch, cluster) ridge penalty.

    _harmony/__init__.py:454-480: harmony1 -> ``ridge_lambda`` everywhere;
    harmony2 -> ``alpha*E`` with the 1e30 sentinel where ``O/N_b < threshold``
    or ``N_b == 0``; always the sentinel where ``O + lambda == 0``.
    """
    dt = E.dtype
    sentinel = dt.type(SUPPRESS_PENALTY)
    if not dynamic_lambda:
        lam = np.full_like(E, ridge_lambda)
    else:
        lam = (alpha * E).astype(dt)
        if threshold is not None:
            N_b = np.asarray(N_b, dtype=dt)
            safe = np.where(N_b > 0, N_b, np.ones_like(N_b))
            mask = (O / safe[:, None]) < threshold
            mask |= (N_b[:, None] == 0)
            lam[mask] = sentinel
    lam[(O + lam) == 0] = sentinel
    return lam


def inv_mats(O, lambda_kb):
    """Closed-form inverse of ``[[N_k, O_k^T], [O_k, diag(O_k + lambda_k)]]``.

    _cuda/harmony/correction/kernels_correction_fast.cuh:6-123 (N_k taken
    from O).  Returns K x (B+1) x (B+1).
    """
    dt = O.dtype.type
    B, K = O.shape
    f = dt(1) / (O + lambda_kb)
    P = -f * O
    Nk = O.sum(axis=0)
    c = Nk - (f * O * O).sum(axis=0)
    cinv = dt(1) / c
    out = np.empty((K, B + 1, B + 1), dtype=O.dtype)
    for k in range(K):
        p = P[:, k]
        out[k, 0, 0] = cinv[k]
        out[k, 0, 1:] = cinv[k] * p
        out[k, 1:, 0] = p * cinv[k]
        out[k, 1:, 1:] = np.outer(p, p) * cinv[k] + np.diag(f[:, k])
    return out


def correction_batched(X, R, O, lambda_kb, cats, n_batches):
    """Batched ridge correction (also the result of the "fast" method).

    correction_batched.cu:18-130: ``Phi_k = [R_k^T X ; R_{k,b}^T X_b]``
    from the original X; ``W_k = inv_k Phi_k``; ``W_k[0] = 0``;
    ``Z = X - sum_k R_ik W_k[cat_i + 1]`` (kernels_outer.cuh:41-65).
    correction_fast.cu:23-109 applies the same per-cluster terms one cluster
    at a time, so it has the same value.
    """
    dt = X.dtype
    inv = inv_mats(O, lambda_kb)
    K = R.shape[1]
    Phi = np.zeros((K, n_batches + 1, X.shape[1]), dtype=dt)
    Phi[:, 0, :] = R.T @ X
    for b in range(n_batches):
        m = cats == b
        if m.any():
            Phi[:, b + 1, :] = R[m].T @ X[m]
    W = np.einsum("kij,kjd->kid", inv, Phi)
    W[:, 0, :] = 0
    Wc = W[:, cats + 1, :]  # K x N x d
    Z = X - np.einsum("nk,knd->nd", R, Wc)
    return Z.astype(dt), W


def correction_original(X, R, lambda_kb, cats, n_batches):
    """Original per-cluster correction with an explicit inverse.

    _harmony/__init__.py:532-570 with the aggregated matrix of
    kernels_scatter.cuh:28-42: ``A = [[sum s, s^T], [s, diag(s)]] +
    diag(0, lambda_k)`` where ``s`` are the per-batch sums of R[:, k].
    """
    dt = X.dtype
    K = R.shape[1]
    Z = X.copy()
    for k in range(K):
        r = R[:, k]
        s = np.zeros(n_batches, dtype=dt)
        np.add.at(s, cats, r)
        A = np.zeros((n_batches + 1, n_batches + 1), dtype=dt)
        A[0, 0] = s.sum()
        A[0, 1:] = s
        A[1:, 0] = s
        A[1:, 1:] = np.diag(s)
        lam = np.zeros(n_batches + 1, dtype=dt)
        lam[1:] = lambda_kb[:, k]
        inv = np.linalg.inv(A + np.diag(lam))
        Phi = np.zeros((n_batches + 1, X.shape[1]), dtype=dt)
        Phi[0] = X.T @ r
        for b in range(n_batches):
            m = cats == b
            Phi[b + 1] = X[m].T @ r[m]
        W = inv @ Phi
        W[0, :] = 0
        Z -= r[:, None] * W[cats + 1]
    return Z


# --------------------------------------------------------------------------
# K-means initialiser (the build's own algorithm; see module docstring)
# --------------------------------------------------------------------------


def kmeans_centroids(Z_norm, n_clusters: int, random_state: int, max_iter: int = 25):
    """Seeded Lloyd k-means used to produce ``Y0`` when none is given.

    The reference calls cuML ``KMeans(init="k-means||", n_init=1, max_iter=25,
    random_state)`` (_harmony/__init__.py:315-323); cuML is not available, so
    this build defines its own deterministic initialiser: the first
    ``n_clusters`` cells of the PCG order for ``random_state`` are the seeds;
    then ``max_iter`` Lloyd steps (assign to ``argmax z.y - |y|^2/2``, lowest
    index on ties; replace each non-empty cluster by its mean).  Parity with
    cuML is not claimed.
    """
    dt = Z_norm.dtype
    n = Z_norm.shape[0]
    seeds = block_order(n, random_state & _U32)[:n_clusters]
    Y = Z_norm[seeds].astype(dt).copy()
    for _ in range(max_iter):
        score = Z_norm @ Y.T - dt.type(0.5) * np.einsum("kd,kd->k", Y, Y)[None, :]
        lab = np.argmax(score, axis=1)
        sums = _scatter_rows(lab, Z_norm, n_clusters)
        cnt = np.bincount(lab, minlength=n_clusters)
        nz = cnt > 0
        Y[nz] = sums[nz] / cnt[nz, None].astype(dt)
    return Y


# --------------------------------------------------------------------------
# Driver
# --------------------------------------------------------------------------


def harmonize(X, cats, n_batches, *, Y0=None, n_clusters=None,
              max_iter_harmony=10, max_iter_clustering=200, tol_harmony=1e-4,
              tol_clustering=1e-5, ridge_lambda=1.0, sigma=0.1,
              block_proportion=0.05, theta=2.0, tau=0, correction_method="batched",
              random_state=0, stabilized_penalty=True, dynamic_lambda=True,
              alpha=0.2, batch_prune_threshold=1e-5, n_b=None, record=None):
    """Whole Harmony run on the CPU.

    _harmony/__init__.py:41-289: normalise, initialise, then per Harmony
    iteration: clustering with seed ``random_state + i*1000003``, lambda_kb,
    correction from the ORIGINAL X, re-normalise, convergence on the last two
    objectives (:722-734; the init objective counts).

    ``record`` (a dict) receives per-iteration diagnostics when given.
    Returns ``Z_hat``.
    """
    X = np.ascontiguousarray(X)
    dt = X.dtype
    n = X.shape[0]
    cats = np.asarray(cats, dtype=np.int32)
    if n_b is None:
        n_b = np.bincount(cats, minlength=n_batches)
    N_b = np.asarray(n_b, dtype=dt)
    Pr_b = (N_b.reshape(-1, 1) / n).astype(dt)
    if n_clusters is None:
        n_clusters = max(int(min(100, n / 30)), 2)
    th = theta_array(theta, n_batches, dt, tau=tau, n_b=N_b, n_clusters=n_clusters)
    Z_norm = l2_normalize(X)
    if Y0 is None:
        Y0 = kmeans_centroids(Z_norm, n_clusters, random_state)
    R, E, O, obj0 = init_clusters(Z_norm, Y0, sigma=sigma, Pr_b=Pr_b, theta=th,
                                  cats=cats, n_batches=n_batches,
                                  stabilized=stabilized_penalty)
    objectives = [obj0]
    block_size = int(n * block_proportion)
    Z_hat = X
    n_iters = []
    for i in range(max_iter_harmony):
        last, nc, _ = clustering_loop(
            Z_norm, R, E, O, Pr_b, cats, th, sigma=sigma, tol=tol_clustering,
            max_iter=max_iter_clustering, block_size=block_size,
            seed=(random_state + i * 1000003) & _U32,
            stabilized=stabilized_penalty)
        objectives.append(last)
        n_iters.append(nc)
        lam = compute_lambda_kb(E, O, N_b, alpha=alpha,
                                threshold=batch_prune_threshold,
                                ridge_lambda=ridge_lambda,
                                dynamic_lambda=dynamic_lambda)
        if correction_method == "original":
            Z_hat = correction_original(X, R, lam, cats, n_batches)
        else:
            Z_hat, _ = correction_batched(X, R, O, lam, cats, n_batches)
        Z_norm = l2_normalize(Z_hat)
        if len(objectives) >= 2:
            old, new = objectives[-2], objectives[-1]
            if (old - new) < tol_harmony * abs(old):
                break
    if record is not None:
        record.update(objectives=objectives, clustering_iters=n_iters, R=R, E=E,
                      O=O, Z_norm=Z_norm)
    return Z_hat


# --------------------------------------------------------------------------
# Several covariates at once (stacked one-hot Phi).  Build extension: the
# reference has no counterpart -- a list of keys is joined into one
# categorical (_harmony/_helper.py:223-236).  Here each key is a covariate of
# its own, as in the Harmony paper (Korsunsky et al. 2019): Phi stacks one
# one-hot block per covariate, so every cell has one level in each.
#
# * O = Phi^T R (levels x K); E[b] = Pr_b (x) colsum(R), Pr_b = N_b / N per
#   level (each covariate's Pr sums to 1).
# * The R update of a cell multiplies the penalties of its levels:
#   R_ik ~ exp(-2/sigma (1 - sim_ik)) * prod_b pen_bk^(Phi_ib), which is the
#   minimiser of the objective below for fixed O and E (for one covariate it
#   is the reference's update).
# * Objective: k-means + sigma * entropy + sigma * sum over all levels of
#   theta_b O_bk log(numer / (E + 1)).
# * Correction per cluster k: A_k = Phi~^T diag(R_k) Phi~ + diag(0, lambda_k)
#   with Phi~ = [1 | Phi]; W_k = A_k^{-1} Phi~^T diag(R_k) X; W_k[0] = 0;
#   Z_i = X_i - sum_k R_ik sum_{b: Phi_ib = 1} W_k[1 + b].  A_k is no longer
#   an arrowhead matrix (the cross-covariate blocks hold co-occurrence sums),
#   so it is solved densely.
# --------------------------------------------------------------------------


def multi_levels(codes_list, n_levels_list):
    """Global level ids (N x C, covariate c's codes offset by the levels of the
    covariates before it) and the per-level cell counts."""
    offs = np.concatenate([[0], np.cumsum(n_levels_list)[:-1]]).astype(np.int64)
    lev = np.stack([np.asarray(c, np.int64) + o for c, o in zip(codes_list, offs)], axis=1)
    n_b = np.bincount(lev.ravel(), minlength=int(np.sum(n_levels_list)))
    return lev.astype(np.int32), n_b


def multi_theta(theta, n_levels_list, dtype, *, tau=0, n_b=None, n_clusters=None):
    """theta per level: a scalar, or one value per covariate broadcast to its
    levels, or one value per level; then the reference's tau discount
    (_harmony/__init__.py:172-175) per level."""
    dt = np.dtype(dtype)
    total = int(np.sum(n_levels_list))
    if isinstance(theta, (int, float)):
        th = np.full(total, float(theta), dtype=dt)
    else:
        th = np.asarray(theta, dtype=dt).ravel()
        if th.size == len(n_levels_list):
            th = np.repeat(th, n_levels_list).astype(dt)
        elif th.size != total:
            raise ValueError(f"theta needs 1, {len(n_levels_list)} or {total} values")
    if tau > 0:
        th = th * (1 - np.exp(-np.asarray(n_b, dt) / (n_clusters * tau)) ** 2)
    return np.ascontiguousarray(th.astype(dt))


def _phi_rows(lev, rows, n_levels):
    """Phi^T rows: the sum of ``rows`` over the cells of every level."""
    out = np.zeros((n_levels, rows.shape[1]), dtype=rows.dtype)
    for c in range(lev.shape[1]):
        out += _scatter_rows(lev[:, c], rows, n_levels)
    return out


def multi_penalty_rows(pen, lev):
    """Per-cell penalty: the product of the cell's level penalties."""
    out = pen[lev[:, 0]].copy()
    for c in range(1, lev.shape[1]):
        out *= pen[lev[:, c]]
    return out


def multi_init(Z_norm, Y0, *, sigma, Pr_b, theta, lev, n_levels, stabilized):
    dt = Z_norm.dtype
    Y_norm = l2_normalize(np.asarray(Y0, dtype=dt))
    sim = Z_norm @ Y_norm.T
    R = l1_normalize(np.exp(dt.type(-2 / sigma) * (dt.type(1) - sim)))
    E = (Pr_b.reshape(-1, 1) * R.sum(axis=0)[None, :]).astype(dt)
    O = _phi_rows(lev, R, n_levels)
    obj = objective(R, sim, O, E, theta, sigma, stabilized)
    return R, E, O, float(obj)


def multi_clustering_loop(Z_norm, R, E, O, Pr_b, lev, theta, *, sigma, tol, max_iter,
                          block_size, seed, stabilized):
    """clustering_loop with stacked one-hot Phi (same block order, same
    remove -> penalty -> rows -> add order, same window-3 test)."""
    dt = R.dtype.type
    n = R.shape[0]
    L = O.shape[0]
    term = dt(-2) / dt(sigma)
    tol_t = dt(tol)
    Pr = Pr_b.reshape(-1)
    objectives = []
    for it in range(max_iter):
        Y_norm = l2_normalize(R.T @ Z_norm)
        sim = Z_norm @ Y_norm.T
        order = block_order(n, (seed + it) & _U32)
        for pos in range(0, n, block_size):
            idx = order[pos:pos + block_size]
            R_in = R[idx]
            lv = lev[idx]
            O -= _phi_rows(lv, R_in, L)
            E -= Pr[:, None] * R_in.sum(axis=0)[None, :]
            pen = penalty(E, O, theta, stabilized)
            v = np.exp(term * (dt(1) - sim[idx])) * multi_penalty_rows(pen, lv)
            R_out = v * (dt(1) / v.sum(axis=1))[:, None]
            R[idx] = R_out
            O += _phi_rows(lv, R_out, L)
            E += Pr[:, None] * R_out.sum(axis=0)[None, :]
        objectives.append(objective(R, sim, O, E, theta, sigma, stabilized))
        if len(objectives) >= 4:
            old = dt(0)
            new = dt(0)
            for w in range(3):
                old = dt(old + objectives[-2 - w])
                new = dt(new + objectives[-1 - w])
            if (old - new) < tol_t * abs(old):
                break
    last = objectives[-1] if objectives else dt(0)
    return float(last), len(objectives), [float(o) for o in objectives]


def multi_correction(X, R, lambda_kb, lev, n_levels):
    """Ridge correction with stacked one-hot Phi (dense solve per cluster).
    Returns (Z_hat, W) with W of shape K x (levels + 1) x d."""
    dt = X.dtype
    n, d = X.shape
    K = R.shape[1]
    P = np.zeros((n, n_levels + 1), dtype=dt)
    P[:, 0] = 1
    for c in range(lev.shape[1]):
        P[np.arange(n), 1 + lev[:, c]] = 1
    W = np.zeros((K, n_levels + 1, d), dtype=dt)
    Z = X.copy()
    for k in range(K):
        PR = P * R[:, k:k + 1]
        A = PR.T @ P
        A[np.arange(1, n_levels + 1), np.arange(1, n_levels + 1)] += lambda_kb[:, k]
        W[k] = np.linalg.solve(A, PR.T @ X)
        W[k, 0] = 0
        Z -= R[:, k:k + 1] * (P @ W[k])
    return Z.astype(dt), W


def harmonize_multi(X, codes_list, n_levels_list, *, Y0, n_clusters=None, max_iter_harmony=10,
                    max_iter_clustering=200, tol_harmony=1e-4, tol_clustering=1e-5,
                    ridge_lambda=1.0, sigma=0.1, block_proportion=0.05, theta=2.0, tau=0,
                    random_state=0, stabilized_penalty=True, dynamic_lambda=True, alpha=0.2,
                    batch_prune_threshold=1e-5, record=None):
    """Whole Harmony run with one covariate per entry of ``codes_list``
    (stacked one-hot Phi); same driver as ``harmonize`` otherwise."""
    X = np.ascontiguousarray(X)
    dt = X.dtype
    n = X.shape[0]
    lev, n_b = multi_levels(codes_list, n_levels_list)
    L = int(np.sum(n_levels_list))
    N_b = n_b.astype(dt)
    Pr_b = (N_b.reshape(-1, 1) / n).astype(dt)
    if n_clusters is None:
        n_clusters = max(int(min(100, n / 30)), 2)
    th = multi_theta(theta, n_levels_list, dt, tau=tau, n_b=N_b, n_clusters=n_clusters)
    Z_norm = l2_normalize(X)
    R, E, O, obj0 = multi_init(Z_norm, Y0, sigma=sigma, Pr_b=Pr_b, theta=th, lev=lev,
                               n_levels=L, stabilized=stabilized_penalty)
    objectives = [obj0]
    block_size = int(n * block_proportion)
    Z_hat = X
    n_iters = []
    for i in range(max_iter_harmony):
        last, nc, _ = multi_clustering_loop(
            Z_norm, R, E, O, Pr_b, lev, th, sigma=sigma, tol=tol_clustering,
            max_iter=max_iter_clustering, block_size=block_size,
            seed=(random_state + i * 1000003) & _U32, stabilized=stabilized_penalty)
        objectives.append(last)
        n_iters.append(nc)
        lam = compute_lambda_kb(E, O, N_b, alpha=alpha, threshold=batch_prune_threshold,
                                ridge_lambda=ridge_lambda, dynamic_lambda=dynamic_lambda)
        Z_hat, _ = multi_correction(X, R, lam, lev, L)
        Z_norm = l2_normalize(Z_hat)
        if (objectives[-2] - objectives[-1]) < tol_harmony * abs(objectives[-2]):
            break
    if record is not None:
        record.update(objectives=objectives, clustering_iters=n_iters, R=R, E=E, O=O,
                      Z_norm=Z_norm, levels=lev)
    return Z_hat
