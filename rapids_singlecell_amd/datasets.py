"""Synthetic PCA embeddings for tests and benchmarks (Gaussian cell-type
mixtures with per-batch shifts)."""

from __future__ import annotations

import numpy as np
import pandas as pd


def synthetic_pcs(n_cells: int, n_pcs: int = 50, n_types: int = 10, n_batches: int = 3, *,
                  dirichlet_batches: bool = False, center_scale: float = 3.0,
                  shift_scale: float = 1.0, seed: int = 0):
    """Return ``(X float32 [n_cells, n_pcs], obs DataFrame with 'batch' and 'type')``.

    centers ~ N(0, center_scale^2); cell type uniform; batch uniform, or drawn
    with Dirichlet(1) proportions when ``dirichlet_batches``; per-batch shift
    ~ N(0, shift_scale^2); unit isotropic noise.
    """
    rng = np.random.default_rng(seed)
    centers = rng.normal(0.0, center_scale, size=(n_types, n_pcs))
    ctype = rng.integers(n_types, size=n_cells)
    if dirichlet_batches:
        p = rng.dirichlet(np.ones(n_batches))
        batch = rng.choice(n_batches, size=n_cells, p=p)
    else:
        batch = rng.integers(n_batches, size=n_cells)
    shifts = rng.normal(0.0, shift_scale, size=(n_batches, n_pcs))
    X = centers[ctype] + rng.normal(size=(n_cells, n_pcs)) + shifts[batch]
    obs = pd.DataFrame(
        {
            "batch": pd.Categorical([f"b{b:03d}" for b in batch]),
            "type": pd.Categorical([f"t{t:03d}" for t in ctype]),
        }
    )
    return X.astype(np.float32), obs


class SimpleAnnData:
    """Minimal AnnData stand-in: ``obsm`` dict and ``obs`` DataFrame."""

    def __init__(self, obs: pd.DataFrame, obsm: dict | None = None):
        self.obs = obs
        self.obsm = dict(obsm or {})

    @property
    def n_obs(self) -> int:
        return len(self.obs)
