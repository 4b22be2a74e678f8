"""``pp.pca`` for dense data with zero centring: the producer of the Harmony
input ``obsm["X_pca"]``.

Follows preprocessing/_pca.py:71-330 of rapids_singlecell for a dense
matrix and ``zero_center=True``: the same defaults (50 components, or
min(n_obs, n_vars) - 1), ``mask_var`` / ``use_highly_variable`` selection,
result keys (``obsm["X_pca"]``, ``varm["PCs"]``, ``uns["pca"]`` with
``variance`` and ``variance_ratio``) and sign convention (each component's
largest-magnitude loading is positive, as scikit-learn's ``svd_flip`` on the
components).  The column means and the covariance of the centred matrix run
on matrix cores (``hz_pca_covariance``); the symmetric eigendecomposition of
the n_vars x n_vars covariance runs on the host in float64 (its cost does not
depend on the number of cells); the projection runs on matrix cores
(``hz_pca_project``).
"""

from __future__ import annotations

import copy as _copy
import ctypes

import numpy as np
import torch

from .. import _native


def _ptr(t: torch.Tensor) -> ctypes.c_void_p:
    return ctypes.c_void_p(t.data_ptr())


def pca_dense(X, n_comps: int):
    """(X_pca device tensor float32, components n_vars x n_comps float64,
    explained variance, explained variance ratio) of a dense matrix."""
    lib = _native.load()
    Xd = X if isinstance(X, torch.Tensor) else torch.as_tensor(np.ascontiguousarray(X))
    Xd = Xd.to(device="cuda", dtype=torch.float32).contiguous()
    n, g = Xd.shape
    if n < 2:
        raise ValueError("PCA needs at least two observations")
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    nb = lib.hz_pca_workspace_bytes(n, g)
    ws = torch.empty(max(nb, 256), dtype=torch.uint8, device=Xd.device)
    mean = torch.empty(g, dtype=torch.float64, device=Xd.device)
    cov = torch.empty((g, g), dtype=torch.float64, device=Xd.device)
    _native.check(lib.hz_pca_covariance(_ptr(Xd), n, g, _ptr(mean), _ptr(cov), _ptr(ws),
                                        ws.numel(), s), "hz_pca_covariance")
    C = cov.cpu().numpy()
    evals, evecs = np.linalg.eigh(C)
    evals = np.clip(evals[::-1], 0.0, None)
    evecs = evecs[:, ::-1]
    total = float(evals.sum())
    comps = np.ascontiguousarray(evecs[:, :n_comps])
    # deterministic signs: the largest-magnitude loading of each component > 0
    pick = np.argmax(np.abs(comps), axis=0)
    signs = np.sign(comps[pick, np.arange(n_comps)])
    signs[signs == 0] = 1.0
    comps *= signs[None, :]
    V = torch.as_tensor(comps.astype(np.float32), device=Xd.device)
    out = torch.empty((n, n_comps), dtype=torch.float32, device=Xd.device)
    _native.check(lib.hz_pca_project(_ptr(Xd), n, g, _ptr(mean), _ptr(V), n_comps, _ptr(out),
                                     _ptr(ws), ws.numel(), s), "hz_pca_project")
    var = evals[:n_comps].copy()
    ratio = var / total if total > 0 else np.zeros_like(var)
    return out, comps, var, ratio


def pca(adata, n_comps: int | None = None, *, layer: str | None = None, zero_center: bool = True,
        svd_solver: str | None = None, random_state: int | None = 0, mask_var=None,
        use_highly_variable: bool | None = None, dtype: str = "float32", chunked: bool = False,
        chunk_size: int | None = None, key_added: str | None = None, copy: bool = False,
        **kwargs):
    """Principal component analysis of a dense ``adata.X`` (or ``layer``)
    with zero centring (see module docstring)."""
    if use_highly_variable is True and "highly_variable" not in adata.var.keys():
        raise ValueError(
            "Did not find adata.var['highly_variable']. "
            "Either your data already only consists of highly-variable genes "
            "or consider running `highly_variable_genes` first."
        )
    if not zero_center:
        raise NotImplementedError("zero_center=False (truncated SVD) is not provided")
    if chunked:
        raise NotImplementedError("chunked (incremental) PCA is not provided")
    adata = _copy.deepcopy(adata) if copy else adata
    X = adata.layers[layer] if layer is not None else adata.X
    if hasattr(X, "toarray") and not isinstance(X, (np.ndarray, torch.Tensor)):
        raise NotImplementedError("sparse input is not provided: pass a dense matrix")
    mask_param = None
    mask = None
    if use_highly_variable:
        mask_param = "highly_variable"
        mask = np.asarray(adata.var["highly_variable"].values, dtype=bool)
    elif isinstance(mask_var, str):
        mask_param = mask_var
        mask = np.asarray(adata.var[mask_var].values, dtype=bool)
    elif mask_var is not None:
        mask = np.asarray(mask_var, dtype=bool)
    if mask is not None:
        X = X[:, mask] if not isinstance(X, torch.Tensor) else X[:, torch.as_tensor(mask)]
    if n_comps is None:
        min_dim = min(X.shape[0], X.shape[1])
        n_comps = min_dim - 1 if 50 >= min_dim else 50
    out, comps, var, ratio = pca_dense(X, n_comps)
    X_pca = out.cpu().numpy().astype(dtype)
    key_obsm, key_varm, key_uns = ("X_pca", "PCs", "pca") if key_added is None else [key_added] * 3
    adata.obsm[key_obsm] = X_pca
    if not hasattr(adata, "uns"):
        adata.uns = {}
    adata.uns[key_uns] = {
        "params": {"zero_center": zero_center, "use_highly_variable": mask_param == "highly_variable",
                   "mask_var": mask_param, **({"layer": layer} if layer is not None else {})},
        "variance": var,
        "variance_ratio": ratio,
    }
    if not hasattr(adata, "varm"):
        adata.varm = {}
    if mask is not None:
        full = np.zeros((len(mask), n_comps))
        full[mask] = comps
        adata.varm[key_varm] = full
    else:
        adata.varm[key_varm] = comps
    return adata if copy else None
