"""Several covariates at once (stacked one-hot Phi) on the GPU against the
oracle's ``harmonize_multi``.

Each key is a covariate of its own: O and E have one row per level, a cell's
penalty is the product of its levels' penalties, and the ridge correction
solves the full normal equations of [1 | Phi] per cluster.  Bounds (relative
Frobenius error of Z_corr): FP64 <= 1e-8 with identical clustering iteration
counts; plain FP32 <= 1e-4 and split-BF16 <= 1e-3 against the FP64 oracle.
A covariate given twice must reproduce the one-covariate GPU run with theta
doubled and lambda halved (the reduction the oracle is pinned by).
"""

from __future__ import annotations

import numpy as np
import pandas as pd
import pytest
import torch

from oracle import harmony_oracle as ho
from rapids_singlecell_amd import _native
from rapids_singlecell_amd.preprocessing._harmony import harmonize

pytestmark = pytest.mark.gpu

FIXED = dict(tol_harmony=float("-inf"), tol_clustering=float("-inf"))


def _data(n, d, B, D, K, seed):
    rng = np.random.default_rng(seed)
    centers = rng.normal(0, 3, size=(8, d))
    b = rng.integers(0, B, n)
    dn = rng.integers(0, D, n)
    X = (centers[rng.integers(0, 8, n)] + rng.normal(size=(n, d))
         + rng.normal(0, 0.5, size=(B, d))[b] + rng.normal(0, 0.5, size=(D, d))[dn])
    obs = pd.DataFrame({"batch": pd.Categorical([f"b{v:02d}" for v in b]),
                        "donor": pd.Categorical([f"d{v:02d}" for v in dn])})
    Z = ho.l2_normalize(X)
    Y0 = Z[rng.choice(n, K, replace=False)]
    return X, obs, b.astype(np.int32), dn.astype(np.int32), Y0


def _gpu(X, obs, keys, Y0, dtype, **kw):
    rec = {}
    Z = harmonize(torch.as_tensor(X.astype(dtype), device="cuda"), obs, keys,
                  init_centroids=Y0.astype(dtype), record=rec, **kw)
    return Z.cpu().numpy().astype(np.float64), rec


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("stabilized,dynamic", [(False, False), (True, True)])
def test_multi_fp64_matches_oracle(device, stabilized, dynamic):
    X, obs, b, dn, Y0 = _data(3000, 20, 3, 5, 10, seed=1)
    kw = dict(n_clusters=10, max_iter_harmony=3, max_iter_clustering=8, random_state=4,
              stabilized_penalty=stabilized, dynamic_lambda=dynamic)
    rr = {}
    Zr = ho.harmonize_multi(X, [b, dn], [3, 5], Y0=Y0, record=rr, **kw)
    Zg, rg = _gpu(X, obs, ["batch", "donor"], Y0, np.float64, multi_covariate=True, **kw)
    e = _rel(Zg, Zr)
    print(f"\nmulti-covariate fp64 (3 x 5 levels): clustering iters gpu={rg['clustering_iters']} "
          f"oracle={rr['clustering_iters']}, rel err Z_corr = {e:.3e}", flush=True)
    assert rg["clustering_iters"] == rr["clustering_iters"]
    np.testing.assert_allclose(rg["objectives"], rr["objectives"], rtol=1e-9)
    np.testing.assert_allclose(rg["O"].cpu().numpy(), rr["O"], rtol=1e-8, atol=1e-8)
    assert e < 1e-8


@pytest.mark.parametrize("precision,bound", [("native", 1e-4), ("split_bf16", 1e-3)])
def test_multi_fp32_matches_oracle(device, precision, bound):
    X, obs, b, dn, Y0 = _data(12000, 30, 4, 12, 20, seed=2)
    kw = dict(n_clusters=20, max_iter_harmony=2, max_iter_clustering=10, random_state=1,
              stabilized_penalty=False, dynamic_lambda=False, **FIXED)
    Zr = ho.harmonize_multi(X, [b, dn], [4, 12], Y0=Y0, **kw)
    Zg, _ = _gpu(X, obs, ["batch", "donor"], Y0, np.float32, multi_covariate=True,
                 precision=precision, **kw)
    e = _rel(Zg, Zr)
    print(f"\nmulti-covariate fp32 {precision} (4 x 12 levels): rel err Z_corr vs FP64 oracle = "
          f"{e:.3e}", flush=True)
    assert e < bound


def test_multi_c3_shape_reduced(device):
    """The C3 covariates (16 batches x 64 donors, 80 levels) at 40k cells."""
    X, obs, b, dn, Y0 = _data(40000, 50, 16, 64, 40, seed=3)
    kw = dict(n_clusters=40, max_iter_harmony=2, max_iter_clustering=10, random_state=0,
              stabilized_penalty=False, dynamic_lambda=False, ridge_lambda=1.0, **FIXED)
    Zr = ho.harmonize_multi(X, [b, dn], [16, 64], Y0=Y0, **kw)
    Z64, _ = _gpu(X, obs, ["batch", "donor"], Y0, np.float64, multi_covariate=True, **kw)
    Z32, _ = _gpu(X, obs, ["batch", "donor"], Y0, np.float32, multi_covariate=True, **kw)
    e64, e32 = _rel(Z64, Zr), _rel(Z32, Zr)
    print(f"\nmulti-covariate 16 x 64 levels, 40k cells: rel err Z_corr fp64 = {e64:.3e}, "
          f"fp32 = {e32:.3e}", flush=True)
    assert e64 < 1e-8 and e32 < 1e-4


def test_multi_duplicated_covariate_equals_one_covariate(device):
    X, obs, _, _, Y0 = _data(5000, 16, 4, 3, 12, seed=5)
    obs = obs.assign(batch2=obs["batch"])
    kw = dict(n_clusters=12, max_iter_harmony=2, max_iter_clustering=6, random_state=2,
              stabilized_penalty=False, dynamic_lambda=False, **FIXED)
    Z1, r1 = _gpu(X, obs, "batch", Y0, np.float64, theta=3.0, ridge_lambda=0.5, **kw)
    Z2, r2 = _gpu(X, obs, ["batch", "batch2"], Y0, np.float64, multi_covariate=True, theta=1.5,
                  ridge_lambda=1.0, **kw)
    e = _rel(Z2, Z1)
    print(f"\nduplicated covariate vs one covariate (2 theta, lambda / 2): rel err = {e:.3e}",
          flush=True)
    np.testing.assert_allclose(r2["objectives"], r1["objectives"], rtol=1e-10)
    assert e < 1e-9


@pytest.mark.parametrize("dtype,precision", [(np.float64, "native"), (np.float32, "native"),
                                             (np.float32, "split_bf16")])
def test_multi_update_launch_is_bitwise_the_same(device, dtype, precision):
    """80 levels x 60 clusters (4800 entries): the O/E/penalty update between
    blocks runs in a launch of its own spread over the clusters, the automatic
    choice above 512 entries.  Forcing it into the row launch's last
    workgroup gives the same bits; the FP64 run also matches the oracle."""
    X, obs, b, dn, Y0 = _data(20000, 20, 16, 64, 60, seed=6)
    kw = dict(n_clusters=60, max_iter_harmony=2, max_iter_clustering=6, random_state=3,
              stabilized_penalty=False, dynamic_lambda=False, ridge_lambda=1.0, **FIXED)
    before = _native.get_tuning()
    res = {}
    try:
        for mode in (0, 1, 2):  # automatic, last workgroup, own launch
            _native.set_tuning(block_update=mode)
            res[mode] = _gpu(X, obs, ["batch", "donor"], Y0, dtype, multi_covariate=True,
                             precision=precision, **kw)
    finally:
        _native.set_tuning(**before)
    np.testing.assert_array_equal(res[2][0], res[1][0])
    np.testing.assert_array_equal(res[0][0], res[2][0])
    np.testing.assert_array_equal(np.asarray(res[2][1]["objectives"]),
                                  np.asarray(res[1][1]["objectives"]))
    if dtype == np.float64:
        Zr = ho.harmonize_multi(X, [b, dn], [16, 64], Y0=Y0, **kw)
        e = _rel(res[2][0], Zr)
        print(f"\nmulti-covariate fp64, 80 levels x 60 clusters, update in its own launch: "
              f"rel err Z_corr vs oracle = {e:.3e}", flush=True)
        assert e < 1e-8
