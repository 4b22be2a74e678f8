"""Batched form: several Harmony instances advanced together (GPU).

``harmonize_many`` runs the clustering step of instances that share a kernel
shape through one ``hz_clustering_loop_batched`` call (one cooperative sweep
launch per clustering iteration for all of them).  Every block sum and
objective term is fixed point, so each instance's Z_hat, objectives and
iteration counts must be bitwise those of ``harmonize`` on it alone.  One
instance is also checked against the FP64 oracle.
"""

from __future__ import annotations

import ctypes

import numpy as np
import pytest
import torch

from oracle import harmony_oracle as ho
from rapids_singlecell_amd import _native
from rapids_singlecell_amd.datasets import synthetic_pcs
from rapids_singlecell_amd.preprocessing._harmony import harmonize
from rapids_singlecell_amd.preprocessing._harmony._batch import harmonize_many

pytestmark = pytest.mark.gpu


def _problem(n, d, B, K, seed, precision="native", dtype=np.float32, **kw):
    X, obs = synthetic_pcs(n, d, 8, B, seed=seed)
    Z = ho.l2_normalize(X.astype(np.float64))
    Y0 = Z[np.random.default_rng(seed).choice(n, K, replace=False)].astype(dtype)
    p = dict(X=torch.as_tensor(X.astype(dtype), device="cuda"), obs=obs, key="batch",
             n_clusters=K, init_centroids=Y0, precision=precision, random_state=seed,
             max_iter_harmony=3, max_iter_clustering=8)
    p.update(kw)
    return p


def _alone(p):
    kw = {k: v for k, v in p.items() if k not in ("X", "obs", "key")}
    rec = {}
    Z = harmonize(p["X"], p["obs"], p["key"], record=rec, **kw)
    return Z.cpu().numpy(), rec


@pytest.mark.parametrize("precision", ["native", "split_bf16"])
def test_batched_equals_alone(device, precision):
    # one sweep class (d = 50, K in 65..112), uneven sizes and batch counts,
    # reference convergence rules (some instances stop early)
    probs = [_problem(n, 50, B, K, seed, precision)
             for n, B, K, seed in [(60_000, 6, 100, 1), (9_000, 3, 80, 2), (25_000, 12, 90, 3),
                                   (4_000, 2, 70, 4)]]
    rec = []
    got = harmonize_many(probs, record=rec)
    for p, z, r in zip(probs, got, rec):
        za, ra = _alone(p)
        assert r["clustering_iters"] == ra["clustering_iters"]
        assert r["objectives"] == pytest.approx(ra["objectives"], rel=0, abs=0)
        np.testing.assert_array_equal(z.cpu().numpy(), za)
    print(f"\nbatched {precision}: clustering iterations {[r['clustering_iters'] for r in rec]}",
          flush=True)


def test_batched_mixed_classes_and_fallback(device):
    probs = [_problem(20_000, 30, 4, 60, 11), _problem(30_000, 50, 5, 100, 12),
             _problem(15_000, 30, 3, 50, 13), _problem(8_000, 20, 3, 40, 14, dtype=np.float64),
             _problem(12_000, 50, 4, 100, 15, precision="split_bf16"),
             # 32 batches x 100 clusters at 100 PCs: the sweep's LDS does not fit
             _problem(20_000, 100, 32, 100, 16)]
    for p in probs:  # fixed iteration counts
        p.update(tol_harmony=float("-inf"), tol_clustering=float("-inf"))
    got = harmonize_many(probs)
    for p, z in zip(probs, got):
        za, _ = _alone(p)
        np.testing.assert_array_equal(z.cpu().numpy(), za)


def test_batched_instance_matches_oracle(device):
    p = _problem(30_000, 40, 5, 60, 21, max_iter_harmony=2, max_iter_clustering=10,
                 tol_harmony=float("-inf"), tol_clustering=float("-inf"), dynamic_lambda=False,
                 stabilized_penalty=False)
    q = _problem(50_000, 40, 7, 60, 22, max_iter_harmony=2, max_iter_clustering=10,
                 tol_harmony=float("-inf"), tol_clustering=float("-inf"))
    got = harmonize_many([p, q])
    X = p["X"].cpu().numpy().astype(np.float64)
    codes, cats, _ = ho.batch_codes(p["obs"], "batch")
    Zr = ho.harmonize(X, codes, len(cats), Y0=p["init_centroids"].astype(np.float64),
                      n_clusters=60, max_iter_harmony=2, max_iter_clustering=10,
                      tol_harmony=float("-inf"), tol_clustering=float("-inf"),
                      dynamic_lambda=False, stabilized_penalty=False, random_state=21)
    z = got[0].cpu().numpy().astype(np.float64)
    err = np.linalg.norm(z - Zr) / np.linalg.norm(Zr)
    print(f"\nbatched instance vs FP64 oracle: rel err Z_corr = {err:.3e}", flush=True)
    assert err < 1e-4


def test_batched_rejects_mixed_shapes(device):
    """The C entry point refuses jobs of different sweep classes (nothing is
    launched) with HZ_ERR_UNSUPPORTED."""
    from rapids_singlecell_amd.preprocessing._harmony._batch import HzJob
    from rapids_singlecell_amd.preprocessing._harmony._job import HarmonyJob

    jobs = []
    for d in (30, 100):
        X, obs = synthetic_pcs(5000, d, 5, 3, seed=d)
        codes = np.asarray(obs["batch"].cat.codes.values, dtype=np.int32)
        j = HarmonyJob(torch.as_tensor(X, device="cuda"), codes, 3, n_clusters=40,
                       theta_np=np.full(3, 2.0, np.float32),
                       N_b_np=np.bincount(codes, minlength=3).astype(np.float32),
                       max_iter_clustering=4)
        jobs.append(j)
    arr = (HzJob * 2)()
    for i, j in enumerate(jobs):
        arr[i].dims = j.dims
        arr[i].max_iter = 4
        arr[i].ws = j.ws.data_ptr()
        arr[i].ws_bytes = j.ws_bytes
        arr[i].Z_norm = j.Z_norm.data_ptr()
        arr[i].R = j.R.data_ptr()
    lib = _native.load()
    rc = lib.hz_clustering_loop_batched(ctypes.cast(arr, ctypes.c_void_p), 2, None, 0,
                                        ctypes.c_void_p(0))
    assert rc == -4
    assert b"kernel shape" in lib.hz_last_error()


def test_batched_eligibility(device):
    """``hz_batched_eligible``: a job whose block sweep needs more LDS than a
    CU has (32 batches x 100 clusters at 100 PCs) runs alone; float32 jobs of
    smaller shapes are eligible."""
    from rapids_singlecell_amd.preprocessing._harmony._batch import sweep_class
    from rapids_singlecell_amd.preprocessing._harmony._job import HarmonyJob

    def job(d, B, K):
        X, obs = synthetic_pcs(20_000, d, 5, B, seed=d + B)
        codes = np.asarray(obs["batch"].cat.codes.values, dtype=np.int32)
        return HarmonyJob(torch.as_tensor(X, device="cuda"), codes, B, n_clusters=K,
                          theta_np=np.full(B, 2.0, np.float32),
                          N_b_np=np.bincount(codes, minlength=B).astype(np.float32),
                          max_iter_clustering=4)

    assert sweep_class(job(100, 32, 100)) is None
    assert sweep_class(job(50, 5, 100)) is not None
    assert sweep_class(job(30, 32, 40)) is not None
