"""Device-resident state of one Harmony run.

``HarmonyJob`` owns the device buffers of a job (X, Z_norm, Z_hat, R, E, O,
workspace) and advances it one Harmony iteration at a time on its own HIP
stream.  ``harmonize`` drives one job to convergence; the benchmark and the
batched/served forms drive many.
"""

from __future__ import annotations

import ctypes

import numpy as np
import torch

from ... import _native

_U32 = 0xFFFFFFFF


def _ptr(t: torch.Tensor) -> ctypes.c_void_p:
    return ctypes.c_void_p(t.data_ptr())


class HarmonyJob:
    def __init__(self, X: torch.Tensor, cats_np: np.ndarray, n_batches: int, *, n_clusters: int,
                 theta_np: np.ndarray, N_b_np: np.ndarray, sigma: float = 0.1,
                 block_proportion: float = 0.05, max_iter_clustering: int = 200,
                 tol_clustering: float = 1e-5, ridge_lambda: float = 1.0,
                 stabilized_penalty: bool = True, dynamic_lambda: bool = True,
                 alpha: float = 0.2, batch_prune_threshold: float | None = 1e-5,
                 correction_method: str = "batched", random_state: int = 0,
                 precision: int = 0, stream: torch.cuda.Stream | None = None,
                 group_levels: np.ndarray | None = None, n_levels0: int = 0):
        """``cats_np`` holds each cell's batch, or with several covariates
        (``group_levels``: n_groups x C level ids) each cell's group; theta and
        N_b then have one entry per level."""
        if not X.is_cuda:
            raise TypeError("HarmonyJob needs X on a HIP device")
        self.lib = _native.load()
        self.dev = X.device
        self.stream = stream if stream is not None else torch.cuda.current_stream(self.dev)
        self.X = X.contiguous()
        self.dtype = X.dtype
        self.dsz = X.element_size()
        self.n_cells, self.n_pcs = X.shape
        self.n_batches = int(n_batches)
        self.n_clusters = int(n_clusters)
        self.sigma = float(sigma)
        self.block_size = int(self.n_cells * block_proportion)
        if self.block_size < 1:
            raise ValueError(
                f"block_proportion={block_proportion} gives an empty block for "
                f"{self.n_cells} cells"
            )
        self.max_iter_clustering = int(max_iter_clustering)
        self.tol_clustering = float(tol_clustering)
        self.ridge_lambda = float(ridge_lambda)
        self.stabilized = int(bool(stabilized_penalty))
        self.dynamic = int(bool(dynamic_lambda))
        self.alpha = float(alpha)
        self.has_thr = int(batch_prune_threshold is not None)
        self.thr = float(batch_prune_threshold) if self.has_thr else 0.0
        self.method = _native.HZ_CORR[correction_method]
        self.random_state = int(random_state)
        self.precision = int(precision)

        if group_levels is not None and np.asarray(group_levels).shape[1] >= 2:
            gl = np.ascontiguousarray(group_levels, dtype=np.int32)
            self.n_covariates = int(gl.shape[1])
            self.n_levels = int(np.asarray(N_b_np).size)
            self.n_levels0 = int(n_levels0)
        else:
            gl = None
            self.n_covariates = 1
            self.n_levels = self.n_batches
            self.n_levels0 = self.n_batches
        counts = np.bincount(cats_np, minlength=self.n_batches).astype(np.int32)
        offsets = np.zeros(self.n_batches + 1, dtype=np.int32)
        np.cumsum(counts, out=offsets[1:])
        cell_indices = np.argsort(cats_np, kind="stable").astype(np.int32)
        np_dtype = np.float32 if self.dtype == torch.float32 else np.float64
        Pr_b = (np.asarray(N_b_np, dtype=np_dtype).reshape(-1, 1) / self.n_cells).astype(np_dtype)

        dev, tdt = self.dev, self.dtype
        with torch.cuda.stream(self.stream):
            self.cats = torch.as_tensor(np.ascontiguousarray(cats_np, dtype=np.int32), device=dev)
            self.cat_offsets = torch.as_tensor(offsets, device=dev)
            self.cell_indices = torch.as_tensor(cell_indices, device=dev)
            self.N_b = torch.as_tensor(np.asarray(N_b_np, dtype=np_dtype), device=dev)
            self.Pr_b = torch.as_tensor(Pr_b.ravel(), device=dev)
            self.theta = torch.as_tensor(np.ascontiguousarray(theta_np, dtype=np_dtype), device=dev)
            self.group_levels = (torch.as_tensor(gl, device=dev) if gl is not None else None)
            self.dims = _native.dims(
                self.n_cells, self.n_pcs, self.n_clusters, self.n_batches, self.block_size,
                self.dsz, self.precision, self.n_covariates,
                self.n_levels if gl is not None else 0, self.n_levels0 if gl is not None else 0,
                self.group_levels.data_ptr() if gl is not None else None)
            self.ws_bytes = self.lib.hz_workspace_bytes(ctypes.byref(self.dims))
            if self.ws_bytes == 0:
                raise ValueError("unsupported Harmony shape: " +
                                 self.lib.hz_last_error().decode())
            self.ws = torch.empty(self.ws_bytes, dtype=torch.uint8, device=dev)
            self.Z_norm = torch.empty_like(self.X)
            self.Z_hat = torch.empty_like(self.X)
            self.R = torch.empty((self.n_cells, self.n_clusters), dtype=tdt, device=dev)
            L = self.n_levels
            self.E = torch.empty((L, self.n_clusters), dtype=tdt, device=dev)
            self.O = torch.empty((L, self.n_clusters), dtype=tdt, device=dev)
            self.lam = torch.empty((L, self.n_clusters), dtype=tdt, device=dev)
            self.obj = torch.zeros(1, dtype=tdt, device=dev)
            self.last_obj = torch.zeros(1, dtype=tdt, device=dev)
        self._n_iter = ctypes.c_int(0)
        self.Y0 = None

    # ------------------------------------------------------------------ helpers
    @property
    def _s(self) -> ctypes.c_void_p:
        return ctypes.c_void_p(self.stream.cuda_stream)

    @staticmethod
    def footprint_bytes(n_cells, n_pcs, n_clusters, n_batches, block_size, dsz, precision=0):
        """Device bytes a job with one covariate needs (workspace + resident arrays)."""
        lib = _native.load()
        dm = _native.dims(n_cells, n_pcs, n_clusters, n_batches, block_size, dsz, precision)
        ws = lib.hz_workspace_bytes(ctypes.byref(dm))
        if ws == 0:
            return 0
        arrays = dsz * n_cells * (3 * n_pcs + n_clusters) + 4 * 3 * n_cells
        return int(ws + arrays + dsz * n_batches * n_clusters * 4)

    # ------------------------------------------------------------------ stages
    def normalize_input(self) -> None:
        rc = self.lib.hz_l2_normalize(self.dsz, _ptr(self.X), _ptr(self.Z_norm), self.n_cells,
                                      self.n_pcs, self._s)
        _native.check(rc, "hz_l2_normalize")

    def kmeans(self, max_iter: int = 25) -> torch.Tensor:
        nbytes = self.lib.hz_kmeans_workspace_bytes(self.dsz, self.n_cells, self.n_pcs,
                                                    self.n_clusters)
        if nbytes == 0:
            _native.check(-4, "hz_kmeans_workspace_bytes")
        with torch.cuda.stream(self.stream):
            ws = torch.empty(nbytes, dtype=torch.uint8, device=self.dev)
            Y = torch.empty((self.n_clusters, self.n_pcs), dtype=self.dtype, device=self.dev)
        rc = self.lib.hz_kmeans(self.dsz, _ptr(self.Z_norm), self.n_cells, self.n_pcs,
                                self.n_clusters, self.random_state & _U32, max_iter, _ptr(Y),
                                _ptr(ws), nbytes, self._s)
        _native.check(rc, "hz_kmeans")
        ws.record_stream(self.stream)
        return Y

    def init(self, Y0: torch.Tensor | None = None) -> None:
        """Normalise X, pick centroids (k-means unless given), initialise R/E/O.
        The init objective is left in ``self.obj`` (device)."""
        self.normalize_input()
        if Y0 is None:
            Y0 = self.kmeans()
        self.Y0 = Y0.to(device=self.dev, dtype=self.dtype).contiguous()
        rc = self.lib.hz_init_clusters(ctypes.byref(self.dims), _ptr(self.Z_norm), _ptr(self.Y0),
                                       _ptr(self.cats), _ptr(self.cat_offsets),
                                       _ptr(self.cell_indices), _ptr(self.Pr_b), _ptr(self.theta),
                                       self.sigma, self.stabilized, _ptr(self.R), _ptr(self.E),
                                       _ptr(self.O), _ptr(self.obj), _ptr(self.ws), self.ws_bytes,
                                       self._s)
        _native.check(rc, "hz_init_clusters")

    def clustering(self, harmony_iter: int) -> int:
        seed = (self.random_state + harmony_iter * 1000003) & _U32
        rc = self.lib.hz_clustering_loop(ctypes.byref(self.dims), _ptr(self.Z_norm), _ptr(self.R),
                                         _ptr(self.E), _ptr(self.O), _ptr(self.Pr_b),
                                         _ptr(self.cats), _ptr(self.theta), self.sigma,
                                         self.tol_clustering, self.max_iter_clustering, seed,
                                         self.stabilized, _ptr(self.last_obj),
                                         ctypes.byref(self._n_iter), _ptr(self.ws), self.ws_bytes,
                                         self._s)
        _native.check(rc, "hz_clustering_loop")
        return self._n_iter.value

    def correction(self) -> None:
        rc = self.lib.hz_correction(ctypes.byref(self.dims), _ptr(self.X), _ptr(self.R),
                                    _ptr(self.O), _ptr(self.E), _ptr(self.N_b),
                                    _ptr(self.cat_offsets), _ptr(self.cell_indices), self.method,
                                    self.dynamic, self.alpha, self.has_thr, self.thr,
                                    self.ridge_lambda, _ptr(self.lam), _ptr(self.Z_hat),
                                    _ptr(self.Z_norm), _ptr(self.ws), self.ws_bytes, self._s)
        _native.check(rc, "hz_correction")

    def step(self, harmony_iter: int) -> int:
        """One Harmony iteration (clustering then correction); returns the number
        of clustering iterations run.  The objective is left in ``last_obj``."""
        n = self.clustering(harmony_iter)
        self.correction()
        return n
