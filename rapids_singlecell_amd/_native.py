"""ctypes binding of ``libharmony_hip.so`` (declared in ``include/harmony_hip.h``).

Plays the role of the reference's lazy nanobind module loader
(``_cuda/__init__.py:16-58``), except that a missing library is an error, not
``None``: there is no CPU fallback for the hot path.
"""

from __future__ import annotations

import atexit
import ctypes
import os
from pathlib import Path

_LIB_PATH = Path(__file__).resolve().parent / "_lib" / "libharmony_hip.so"

HZ_OK = 0
HZ_F32 = 4
HZ_F64 = 8
HZ_PREC_NATIVE = 0
HZ_PREC_SPLIT_BF16 = 1
HZ_CORR = {"batched": 0, "fast": 1, "original": 2}
TUNE_KEYS = {"row_chunk": 0, "centroid_wg": 1, "no_sweep": 2, "sweep_wg": 3, "sweep_plain": 4,
             "spin_limit": 5, "block_update": 6}
PROF_KINDS = ["row", "centroid", "yreduce", "order", "correction", "phi", "apply", "sweep"]

_vp = ctypes.c_void_p
_i = ctypes.c_int
_ll = ctypes.c_longlong
_d = ctypes.c_double
_u32 = ctypes.c_uint32
_sz = ctypes.c_size_t


class HzDims(ctypes.Structure):
    _fields_ = [
        ("n_cells", _i),
        ("n_pcs", _i),
        ("n_clusters", _i),
        ("n_batches", _i),
        ("block_size", _i),
        ("dtype", _i),
        ("precision", _i),
        ("n_covariates", _i),
        ("n_levels", _i),
        ("n_levels0", _i),
        ("group_levels", _vp),
    ]


_P = ctypes.POINTER(HzDims)

# name -> (restype, argtypes)
_SIGNATURES = {
    "hz_version": (_i, []),
    "hz_release_thread": (_i, []),
    "hz_profile_enable": (_i, [_i]),
    "hz_set_tuning": (_i, [_i, _i]),
    "hz_get_tuning": (_i, [_i]),
    "hz_profile_collect": (_i, [ctypes.POINTER(_d), ctypes.POINTER(_ll), _i]),
    "hz_last_error": (ctypes.c_char_p, []),
    "hz_workspace_bytes": (_sz, [_P]),
    "hz_l2_normalize": (_i, [_i, _vp, _vp, _ll, _i, _vp]),
    "hz_block_order_workspace_bytes": (_sz, [_i]),
    "hz_block_order": (_i, [_i, _u32, _vp, _vp, _sz, _vp]),
    "hz_block_order_noalloc": (_i, [_i, _u32, _vp, _vp]),
    "hz_block_membership_workspace_bytes": (_sz, [_i, _i, _i, _i]),
    "hz_block_membership": (_i, [_i, _i, _i, _vp, _u32, _i, _vp, _vp, _vp, _sz, _vp]),
    "hz_init_clusters": (
        _i,
        [_P, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _d, _i, _vp, _vp, _vp, _vp, _vp, _sz, _vp],
    ),
    "hz_clustering_loop": (
        _i,
        [_P, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _d, _d, _i, _u32, _i, _vp,
         ctypes.POINTER(_i), _vp, _sz, _vp],
    ),
    "hz_batched_workspace_bytes": (_sz, [_i, _i]),
    "hz_batched_eligible": (_i, [_vp, _vp]),
    "hz_clustering_loop_batched": (_i, [_vp, _i, _vp, _sz, _vp]),
    "hz_correction": (
        _i,
        [_P, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _i, _i, _d, _i, _d, _d, _vp, _vp, _vp,
         _vp, _sz, _vp],
    ),
    "hz_objective_workspace_bytes": (_sz, []),
    "hz_objective": (
        _i, [_i, _vp, _vp, _vp, _vp, _vp, _d, _i, _i, _i, _i, ctypes.POINTER(_d), _vp, _sz, _vp]
    ),
    "hz_penalty": (_i, [_i, _vp, _vp, _vp, _vp, _i, _i, _i, _vp]),
    "hz_fused_pen_norm": (_i, [_i, _vp, _vp, _vp, _vp, _vp, _d, _i, _i, _vp]),
    "hz_lambda_kb": (_i, [_i, _vp, _vp, _vp, _i, _i, _i, _d, _i, _d, _d, _vp, _vp]),
    "hz_inv_mats": (_i, [_i, _vp, _vp, _vp, _i, _i, _vp]),
    "hz_knn_workspace_bytes": (_sz, [_ll, _i, _i]),
    "hz_knn": (_i, [_i, _vp, _ll, _i, _i, _vp, _vp, _vp, _sz, _vp]),
    "hz_umap_weights": (_i, [_vp, _vp, _ll, _i, _d, _vp, _vp, _vp, _vp]),
    "hz_pca_workspace_bytes": (_sz, [_ll, _i]),
    "hz_pca_covariance": (_i, [_vp, _ll, _i, _vp, _vp, _vp, _sz, _vp]),
    "hz_pca_project": (_i, [_vp, _ll, _i, _vp, _vp, _i, _vp, _vp, _sz, _vp]),
    "hz_kmeans_workspace_bytes": (_sz, [_i, _ll, _i, _i]),
    "hz_kmeans": (_i, [_i, _vp, _ll, _i, _i, _u32, _i, _vp, _vp, _sz, _vp]),
}

_lib = None


def lib_path() -> Path:
    return Path(os.environ.get("HZ_LIBRARY", _LIB_PATH))


def load() -> ctypes.CDLL:
    """Load the HIP library (once).  Raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    path = lib_path()
    if not path.exists():
        raise RuntimeError(
            f"{path} is missing: build it with `python -m rapids_singlecell_amd._build` "
            "(or __graft_entry__.build()); the Harmony path has no CPU fallback."
        )
    lib = ctypes.CDLL(str(path))
    for name, (res, args) in _SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    # the main thread's side stream, events and pinned flags are released
    # before the interpreter (and the HIP runtime under it) shuts down
    atexit.register(release_thread)
    return lib


def release_thread() -> None:
    """Release the library's per-thread device resources of the calling
    thread (side stream, events, pinned flags); recreated on next use."""
    if _lib is not None:
        _lib.hz_release_thread()


def exported_symbols() -> list[str]:
    return list(_SIGNATURES)


class HarmonyNativeError(RuntimeError):
    pass


def check(rc: int, what: str) -> None:
    if rc != HZ_OK:
        msg = load().hz_last_error()
        msg = msg.decode() if msg else ""
        raise HarmonyNativeError(f"{what} failed (status {rc}): {msg}")


def dims(n_cells, n_pcs, n_clusters, n_batches, block_size, dtype_size, precision=0,
         n_covariates=1, n_levels=0, n_levels0=0, group_levels=None):
    """``hz_dims``; for several covariates ``n_batches`` counts the groups of
    cells sharing all levels and ``group_levels`` is a device pointer."""
    return HzDims(int(n_cells), int(n_pcs), int(n_clusters), int(n_batches),
                  int(block_size), int(dtype_size), int(precision), int(n_covariates),
                  int(n_levels), int(n_levels0), group_levels)


def profile_enable(on: bool) -> None:
    load().hz_profile_enable(int(bool(on)))


def profile_collect() -> dict:
    """Summed launch milliseconds and launch counts per kind since the last call."""
    lib = load()
    ms = (_d * len(PROF_KINDS))()
    cnt = (_ll * len(PROF_KINDS))()
    check(lib.hz_profile_collect(ms, cnt, len(PROF_KINDS)), "hz_profile_collect")
    return {k: (ms[i], cnt[i]) for i, k in enumerate(PROF_KINDS)}


def set_tuning(**kw) -> None:
    """Set library tuning knobs (0 = automatic): row_chunk, centroid_wg."""
    lib = load()
    for k, v in kw.items():
        check(lib.hz_set_tuning(TUNE_KEYS[k], int(v)), f"hz_set_tuning({k})")


def get_tuning() -> dict:
    lib = load()
    return {k: lib.hz_get_tuning(i) for k, i in TUNE_KEYS.items()}
