"""Drop-in ``harmony_integrate`` (preprocessing/_harmony_integrate.py:15-250 of
rapids_singlecell): same signature, warnings, errors and result placement."""

from __future__ import annotations

import warnings
from typing import Literal

import numpy as np
import torch


def _torch_dtype(dtype) -> torch.dtype:
    dt = np.dtype(dtype)
    if dt == np.float32:
        return torch.float32
    if dt == np.float64:
        return torch.float64
    raise TypeError(f"dtype must be float32 or float64, got {dt}")


def harmony_integrate(
    adata,
    key: str | list[str],
    *,
    basis: str = "X_pca",
    adjusted_basis: str = "X_pca_harmony",
    dtype: type = np.float64,
    flavor: Literal["harmony2", "harmony1"] = "harmony2",
    n_clusters: int | None = None,
    max_iter_harmony: int = 10,
    max_iter_clustering: int = 200,
    tol_harmony: float = 1e-4,
    tol_clustering: float = 1e-5,
    sigma: float = 0.1,
    theta=2.0,
    tau: int = 0,
    ridge_lambda: float = 1.0,
    alpha: float = 0.2,
    batch_prune_threshold: float | None = 1e-5,
    correction_method: Literal["fast", "original", "batched"] | None = None,
    colsum_algo: str | None = None,
    block_proportion: float = 0.05,
    random_state: int = 0,
    verbose: bool = False,
    precision: str = "native",
    init_centroids=None,
    multi_covariate: bool = False,
) -> None:
    """Integrate different experiments using the Harmony algorithm.

    Runs on the current HIP device (MI355X).  ``adata`` only needs ``.obsm``
    (mapping) and ``.obs`` (pandas DataFrame).  The result is written to
    ``adata.obsm[adjusted_basis]`` as a NumPy array of ``dtype``.

    All parameters have the meaning of ``rapids_singlecell.pp.harmony_integrate``.
    Extensions: ``precision`` ("native" or "split_bf16", the latter for
    float32), ``init_centroids`` (initial cluster centroids, K x d, used
    instead of the built-in seeded k-means) and ``multi_covariate`` (with a
    list of keys: one covariate per key, stacked one-hot Phi, instead of the
    reference's joined categorical).
    """
    from ._harmony import harmonize

    if flavor not in {"harmony1", "harmony2"}:
        raise ValueError(f"flavor must be 'harmony1' or 'harmony2', got {flavor!r}.")
    stabilized_penalty = flavor == "harmony2"
    dynamic_lambda = flavor == "harmony2"

    if flavor == "harmony2" and ridge_lambda != 1.0:
        warnings.warn(
            "ridge_lambda is ignored when flavor='harmony2'; "
            "use alpha to control regularization strength.",
            UserWarning,
            stacklevel=2,
        )
    if flavor == "harmony1":
        if alpha != 0.2:
            warnings.warn(
                "alpha is ignored when flavor='harmony1'; use ridge_lambda instead.",
                UserWarning,
                stacklevel=2,
            )
        if batch_prune_threshold != 1e-5:
            warnings.warn(
                "batch_prune_threshold is ignored when flavor='harmony1'.",
                UserWarning,
                stacklevel=2,
            )

    if basis not in adata.obsm:
        raise ValueError(
            f"The specified basis '{basis}' is not available in adata.obsm. "
            f"Available bases: {list(adata.obsm.keys())}"
        )

    input_data = adata.obsm[basis]
    device = torch.device("cuda", torch.cuda.current_device())
    try:
        tdt = _torch_dtype(dtype)
        if isinstance(input_data, torch.Tensor):
            X = input_data.to(device=device, dtype=tdt).contiguous()
        else:
            try:
                arr = np.array(input_data, dtype=np.dtype(dtype), order="C")
            except Exception as e:
                raise TypeError(
                    f"Could not convert input of type {type(input_data).__name__} to a "
                    f"device array: {str(e)}"
                ) from e
            try:
                X = torch.as_tensor(arr, device=device)
            except torch.cuda.OutOfMemoryError as e:
                raise MemoryError(
                    "Not enough GPU memory to allocate array. "
                    "Try reducing the dataset size or using a GPU with more memory."
                ) from e
        if bool(torch.isnan(X).any()):
            raise ValueError(
                "Input data contains NaN values. Please handle these before running "
                "harmony_integrate."
            )
    except Exception as e:
        raise RuntimeError(f"Error preparing data for Harmony: {str(e)}") from e

    harmony_out = harmonize(
        X,
        adata.obs,
        key,
        n_clusters=n_clusters,
        max_iter_harmony=max_iter_harmony,
        max_iter_clustering=max_iter_clustering,
        tol_harmony=tol_harmony,
        tol_clustering=tol_clustering,
        ridge_lambda=ridge_lambda,
        sigma=sigma,
        block_proportion=block_proportion,
        theta=theta,
        tau=tau,
        correction_method=correction_method,
        colsum_algo=colsum_algo,
        random_state=random_state,
        stabilized_penalty=stabilized_penalty,
        dynamic_lambda=dynamic_lambda,
        alpha=alpha,
        batch_prune_threshold=batch_prune_threshold,
        verbose=verbose,
        init_centroids=init_centroids,
        precision=precision,
        multi_covariate=multi_covariate,
    )
    adata.obsm[adjusted_basis] = harmony_out.cpu().numpy()
