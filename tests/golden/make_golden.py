"""Writes the small golden fixtures under tests/golden/ from the CPU oracle.

    python tests/golden/make_golden.py

Fixtures (NumPy .npz, no pickles):

* ``block_order.npz``  -- PCG block orders (int32) for a grid of (n, seed),
  including seeds that wrap uint32 and n not a multiple of 64.
* ``harmony_small.npz`` -- a 1,500-cell, 12-PC, 3-batch problem with injected
  initial centroids; FP64 Z_hat, objectives and clustering iteration counts for
  harmony1 and harmony2 (batched correction, 3 Harmony iterations).
* ``lambda_kat.npz``   -- the lambda_kb known-answer inputs/outputs of
  rapids_singlecell's tests/test_harmony.py:342-450.

The GPU tests compare the HIP path against these files; the CPU tests check
that the oracle still reproduces them.
"""

from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parents[1]))

from oracle import harmony_oracle as ho  # noqa: E402

ORDER_GRID = [(1, 0), (2, 7), (63, 1), (64, 0xFFFFFFFF), (1000, 1000003), (4097, 42),
              (65537, 2**31 + 5)]


def small_problem():
    rng = np.random.default_rng(20240601)
    n, d, B, K, types = 1500, 12, 3, 8, 6
    centers = rng.standard_normal((types, d)) * 3
    t = rng.integers(0, types, n)
    cats = rng.choice(B, n, p=[0.5, 0.3, 0.2]).astype(np.int32)
    shifts = rng.standard_normal((B, d))
    X = (centers[t] + rng.standard_normal((n, d)) + shifts[cats]).astype(np.float64)
    Z = ho.l2_normalize(X)
    Y0 = Z[rng.choice(n, K, replace=False)].copy()
    return X, cats, B, K, Y0


def main() -> None:
    orders = {f"n{n}_s{s}": ho.block_order(n, s) for n, s in ORDER_GRID}
    np.savez_compressed(HERE / "block_order.npz", **orders)

    X, cats, B, K, Y0 = small_problem()
    out = dict(X=X, cats=cats, Y0=Y0, n_batches=np.int64(B))
    for flavor, stab in (("harmony1", False), ("harmony2", True)):
        rec = {}
        Zh = ho.harmonize(X, cats, B, Y0=Y0, n_clusters=K, max_iter_harmony=3,
                          max_iter_clustering=10, tol_harmony=float("-inf"),
                          tol_clustering=1e-5, stabilized_penalty=stab,
                          dynamic_lambda=stab, random_state=7, record=rec)
        out[f"{flavor}_Z_hat"] = Zh
        out[f"{flavor}_objectives"] = np.asarray(rec["objectives"])
        out[f"{flavor}_clustering_iters"] = np.asarray(rec["clustering_iters"], np.int64)
    np.savez_compressed(HERE / "harmony_small.npz", **out)

    # tests/test_harmony.py:342-383 (pruning), :406-426 (ridge 0), :429-450 (zero denominator)
    kat = dict(
        prune_E=np.ones((4, 3)) * 10,
        prune_O=np.array([[0, 0, 0], [30, 40, 30], [0, 0, 1], [20, 15, 15]], float),
        prune_Nb=np.array([0, 100, 1, 50], float),
        ridge0_E=np.array([[5.0, 0.0]]), ridge0_O=np.array([[10.0, 0.0]]),
        ridge0_Nb=np.array([100.0]),
        zden_E=np.array([[0.0, 5.0]]), zden_O=np.array([[0.0, 10.0]]), zden_Nb=np.array([100.0]),
    )
    kat["prune_out"] = ho.compute_lambda_kb(kat["prune_E"], kat["prune_O"], kat["prune_Nb"],
                                            alpha=0.2, threshold=1e-5, ridge_lambda=1.0,
                                            dynamic_lambda=True)
    kat["ridge0_out"] = ho.compute_lambda_kb(kat["ridge0_E"], kat["ridge0_O"], kat["ridge0_Nb"],
                                             alpha=0.2, threshold=None, ridge_lambda=0.0,
                                             dynamic_lambda=False)
    kat["zden_out"] = ho.compute_lambda_kb(kat["zden_E"], kat["zden_O"], kat["zden_Nb"],
                                           alpha=0.2, threshold=None, ridge_lambda=1.0,
                                           dynamic_lambda=True)
    np.savez_compressed(HERE / "lambda_kat.npz", **kat)
    for f in sorted(HERE.glob("*.npz")):
        print(f.name, f.stat().st_size)


if __name__ == "__main__":
    main()
