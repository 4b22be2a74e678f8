"""Host-side helpers of the Harmony driver (batch codes, theta, CSR maps)."""

from __future__ import annotations

import numpy as np


def _get_batch_codes(batch_mat, batch_key):
    """Categorical batch vector.

    Same semantics as preprocessing/_harmony/_helper.py:223-236 of the
    reference: one key is used as-is; several keys are joined row-wise with
    "," after ``astype(str)``; the result is ``astype("category")``.
    """
    if isinstance(batch_key, str):
        batch_vec = batch_mat[batch_key]
    elif len(batch_key) == 1:
        batch_vec = batch_mat[batch_key[0]]
    else:
        df = batch_mat[list(batch_key)].astype("str")
        batch_vec = df.apply(lambda row: ",".join(row), axis=1)
    return batch_vec.astype("category")


def _create_category_index_mapping(cats: np.ndarray, n_batches: int):
    """``cat_offsets`` (exclusive cumsum of batch counts) and ``cell_indices``
    (cells in batch order, ascending index within a batch).

    Same result as _utils/_multi_gpu.py:202-232 of the reference
    (``lexsort((idx, cats))``), computed on the host: the codes come from
    pandas on the host anyway and the mapping is integer, bit-exact work.
    """
    counts = np.bincount(cats, minlength=n_batches).astype(np.int32)
    offsets = np.zeros(n_batches + 1, dtype=np.int32)
    np.cumsum(counts, out=offsets[1:])
    cell_indices = np.argsort(cats, kind="stable").astype(np.int32)
    return offsets, cell_indices


def _get_theta_array(theta, n_batches: int, dtype) -> np.ndarray:
    """Theta broadcast to one value per batch.

    preprocessing/_harmony/_helper.py:277-307 of the reference (scalars are
    broadcast; arrays must have ``n_batches`` elements).  ``bool`` is an
    ``int`` there, so ``theta=True`` means 1.0.
    """
    dtype = np.dtype(dtype)
    if isinstance(theta, (float, int)):
        return np.ones(n_batches, dtype=dtype) * float(theta)
    if isinstance(theta, (list, np.ndarray)):
        arr = np.asarray(theta, dtype=dtype)
    elif hasattr(theta, "detach") and hasattr(theta, "cpu"):  # torch tensor
        arr = theta.detach().cpu().numpy().astype(dtype)
    else:
        raise ValueError(
            f"Theta must be float, int, list, numpy array, or cupy array, got {type(theta)}"
        )
    if arr.size != n_batches:
        raise ValueError(
            f"Theta array size ({arr.size}) must match number of batches ({n_batches})"
        )
    return arr.ravel()


def _get_covariates(batch_mat, keys):
    """Several covariates, one per key (stacked one-hot Phi).

    Each key is made categorical on its own (pandas category order, as a
    single key is in ``_get_batch_codes``).  Returns ``levels`` (n_cells x C
    int32: covariate c's code offset by the levels of the covariates before
    it), the number of levels per covariate, the groups of cells that share
    all levels (``group_levels``: n_groups x C in lexicographic order) and each
    cell's group.  No counterpart in the reference, which joins several keys
    into one categorical.
    """
    codes, n_levels = [], []
    for k in keys:
        cat = batch_mat[k].astype("category")
        c = np.asarray(cat.cat.codes.values, dtype=np.int64)
        if c.size and c.min() < 0:
            raise ValueError(f"batch key {k!r} has missing values")
        codes.append(c)
        n_levels.append(int(cat.cat.categories.size))
    offs = np.concatenate([[0], np.cumsum(n_levels)[:-1]])
    levels = np.stack([c + o for c, o in zip(codes, offs)], axis=1).astype(np.int32)
    group_levels, group = np.unique(levels, axis=0, return_inverse=True)
    return (levels, n_levels, np.ascontiguousarray(group_levels, dtype=np.int32),
            np.asarray(group, dtype=np.int32).ravel())


def _get_theta_levels(theta, n_levels, dtype) -> np.ndarray:
    """Theta per level for several covariates: a scalar, one value per
    covariate (broadcast to its levels) or one value per level."""
    dtype = np.dtype(dtype)
    total = int(sum(n_levels))
    if isinstance(theta, (float, int)):
        return np.ones(total, dtype=dtype) * float(theta)
    if hasattr(theta, "detach") and hasattr(theta, "cpu"):
        theta = theta.detach().cpu().numpy()
    arr = np.asarray(theta, dtype=dtype).ravel()
    if arr.size == len(n_levels):
        return np.repeat(arr, n_levels).astype(dtype)
    if arr.size == total:
        return arr
    raise ValueError(
        f"Theta needs one value, one per covariate ({len(n_levels)}) or one per level ({total}), "
        f"got {arr.size}"
    )
