"""Dense PCA on the GPU (``hz_pca_covariance``, ``hz_pca_project``,
``pp.pca``) against scikit-learn's PCA in float64 (GPU).

scikit-learn is the oracle here: its ``covariance_eigh`` solver computes the
same quantities (centred covariance, eigendecomposition, svd_flip on the
components) in float64.  Bounds: covariance within 1e-5 relative (FP32
products of centred values), explained variance within 1e-5, components and
the embedding within 1e-4 (same sign convention).
"""

from __future__ import annotations

import numpy as np
import pandas as pd
import pytest
import torch

from rapids_singlecell_amd.datasets import SimpleAnnData
from rapids_singlecell_amd.preprocessing import _pca

pytestmark = pytest.mark.gpu


def _expr(n, g, seed):
    rng = np.random.default_rng(seed)
    # low-rank structure plus noise, non-zero means (log-expression-like)
    W = rng.normal(size=(8, g))
    Z = rng.normal(size=(n, 8)) * np.linspace(3, 1, 8)
    return (np.log1p(np.abs(Z @ W + rng.normal(scale=0.5, size=(n, g))) + 1.0)).astype(np.float32)


@pytest.mark.parametrize("n,g,nc", [(5000, 200, 30), (3001, 130, 50), (700, 64, 10), (20000, 500, 50)])
def test_pca_matches_sklearn(device, n, g, nc):
    from sklearn.decomposition import PCA

    X = _expr(n, g, seed=n + g)
    out, comps, var, ratio = _pca.pca_dense(torch.as_tensor(X, device=device), nc)
    ref = PCA(n_components=nc, svd_solver="covariance_eigh").fit(X.astype(np.float64))
    np.testing.assert_allclose(var, ref.explained_variance_, rtol=1e-5)
    np.testing.assert_allclose(ratio, ref.explained_variance_ratio_, rtol=1e-5)
    # components: well-separated eigenvalues only (a near-degenerate pair may rotate)
    gaps = np.abs(np.diff(np.concatenate([[np.inf], ref.explained_variance_, [0.0]])))
    sep = np.minimum(gaps[:-1], gaps[1:]) > 1e-3 * ref.explained_variance_[0]
    np.testing.assert_allclose(comps[:, sep], ref.components_.T[:, sep], atol=1e-4)
    emb = ref.transform(X.astype(np.float64))
    np.testing.assert_allclose(out.cpu().numpy()[:, sep], emb[:, sep], atol=1e-4 * np.abs(emb).max())
    print(f"\npca n={n} g={g} nc={nc}: separated components {sep.sum()}", flush=True)


def test_pca_api(device):
    X = _expr(4000, 120, seed=1)
    var = pd.DataFrame({"highly_variable": np.arange(120) % 3 != 0})
    ad = SimpleAnnData(pd.DataFrame(index=range(4000)), {})
    ad.X = X
    ad.var = var
    _pca.pca(ad, n_comps=20, use_highly_variable=True)
    assert ad.obsm["X_pca"].shape == (4000, 20) and ad.obsm["X_pca"].dtype == np.float32
    assert ad.varm["PCs"].shape == (120, 20)
    assert (ad.varm["PCs"][~var["highly_variable"].values] == 0).all()
    assert ad.uns["pca"]["params"]["use_highly_variable"] is True
    assert len(ad.uns["pca"]["variance_ratio"]) == 20
    # default number of components: 50, or min dimension - 1
    ad2 = SimpleAnnData(pd.DataFrame(index=range(30)), {})
    ad2.X = X[:30, :40]
    _pca.pca(ad2)
    assert ad2.obsm["X_pca"].shape == (30, 29)
