"""Served Harmony: a request loop with continuous batching on one GPU.

Integration jobs (a host PCA matrix, batch codes and Harmony parameters) are
queued in arrival order.  A dispatcher admits the job at the head of the queue
as soon as a worker slot is free and its device footprint (resident arrays +
library workspace, ``HarmonyJob.footprint_bytes``) fits in what the memory
budget has left.  Each admitted job runs in its own worker thread on its own
HIP stream (default priority; a fixed pool of ``slots`` workers, each with one
stream for its lifetime), so jobs that arrive while others are on the GPU
start as slots and memory free up, and their kernels overlap on the device.
Every kernel of the path is deterministic (fixed-point atomics, fixed
reduction orders), so a job's result does not depend on what else runs.

Latency is measured per job from ``submit`` to the result being on the host.
"""

from __future__ import annotations

import collections
import queue
import threading
import time
from concurrent.futures import Future

import numpy as np
import pandas as pd
import torch

from . import _native
from .preprocessing._harmony import harmonize
from .preprocessing._harmony._job import HarmonyJob


class HarmonyServer:
    """Request loop over one HIP device.

    Parameters
    ----------
    slots
        Jobs on the device at once (worker threads, one stream each).
    memory_budget
        Device bytes the admitted jobs may hold together; default: 80% of the
        free device memory when the server starts.
    device
        HIP device index.
    """

    def __init__(self, slots: int = 4, memory_budget: int | None = None, device: int = 0):
        if slots < 1:
            raise ValueError("slots must be >= 1")
        self.device = torch.device("cuda", device)
        if memory_budget is None:
            free, _ = torch.cuda.mem_get_info(self.device)
            memory_budget = int(0.8 * free)
        self.memory_budget = int(memory_budget)
        self.slots = int(slots)
        self._queue: collections.deque = collections.deque()
        self._cv = threading.Condition()
        self._in_use = 0
        self._running = 0
        self._closed = False
        self.stats = {"admitted": 0, "rejected": 0, "completed": 0, "failed": 0}
        self._work: queue.Queue = queue.Queue()
        self._workers = [threading.Thread(target=self._worker, name=f"harmony-slot-{i}",
                                          daemon=True) for i in range(self.slots)]
        for w in self._workers:
            w.start()
        self._dispatcher = threading.Thread(target=self._dispatch, name="harmony-dispatch",
                                            daemon=True)
        self._dispatcher.start()

    # ------------------------------------------------------------------ API
    def submit(self, X: np.ndarray, codes: np.ndarray, n_batches: int | None = None,
               **params) -> Future:
        """Queue one integration job.  ``params`` are ``harmonize`` keyword
        arguments (``n_clusters``, ``max_iter_harmony``, ``precision``,
        ``init_centroids`` ...).  The future resolves to ``(Z_hat, info)`` with
        ``Z_hat`` a host array and ``info`` holding the latency (s), the number
        of Harmony iterations and the job's device footprint."""
        X = np.ascontiguousarray(X)
        codes = np.asarray(codes, dtype=np.int32)
        if X.ndim != 2 or codes.shape != (X.shape[0],):
            raise ValueError("X must be n_cells x n_pcs and codes must have n_cells entries")
        B = int(n_batches) if n_batches is not None else int(codes.max()) + 1
        fut: Future = Future()
        need = self._footprint(X, B, params)
        job = {"X": X, "codes": codes, "B": B, "params": params, "future": fut,
               "t_submit": time.perf_counter(), "need": need}
        with self._cv:
            if self._closed:
                raise RuntimeError("server is closed")
            if need <= 0 or need > self.memory_budget:
                self.stats["rejected"] += 1
                fut.set_exception(MemoryError(
                    f"job needs {need} device bytes, budget is {self.memory_budget}"))
                return fut
            self._queue.append(job)
            self._cv.notify_all()
        return fut

    def close(self, wait: bool = True) -> None:
        with self._cv:
            self._closed = True
            self._cv.notify_all()
        if wait:
            self._dispatcher.join()
            with self._cv:
                while self._running:
                    self._cv.wait()
            for _ in self._workers:
                self._work.put(None)
            for w in self._workers:
                w.join()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # ------------------------------------------------------------------ internals
    def _footprint(self, X, B, params) -> int:
        n, d = X.shape
        K = params.get("n_clusters") or max(int(min(100, n / 30)), 2)
        bs = int(n * params.get("block_proportion", 0.05))
        prec = 1 if params.get("precision", "native") == "split_bf16" else 0
        dsz = 8 if X.dtype == np.float64 else 4
        fp = HarmonyJob.footprint_bytes(n, d, K, B, max(bs, 1), dsz, prec)
        return int(fp * 1.1) + (16 << 20)  # + allocator slack and k-means workspace

    def _dispatch(self) -> None:
        while True:
            with self._cv:
                while True:
                    if (self._queue and self._running < self.slots
                            and self._queue[0]["need"] <= self.memory_budget - self._in_use):
                        break
                    if self._closed and not self._queue:
                        return
                    self._cv.wait()
                job = self._queue.popleft()
                self._in_use += job["need"]
                self._running += 1
                self.stats["admitted"] += 1
            self._work.put(job)

    def _worker(self) -> None:
        with torch.cuda.device(self.device):
            stream = torch.cuda.Stream(device=self.device)
            try:
                while True:
                    job = self._work.get()
                    if job is None:
                        return
                    self._run(job, stream)
            finally:
                # the library's per-thread side stream and pinned flags
                _native.release_thread()

    def _run(self, job, stream) -> None:
        fut: Future = job["future"]
        try:
            with torch.cuda.device(self.device):
                with torch.cuda.stream(stream):
                    Xd = torch.as_tensor(job["X"], device=self.device)
                    obs = pd.DataFrame({"batch": pd.Categorical.from_codes(
                        job["codes"], categories=list(range(job["B"])))})
                    rec: dict = {}
                    Z = harmonize(Xd, obs, "batch", record=rec, **job["params"])
                    out = Z.cpu().numpy()
                stream.synchronize()
            info = {"latency_s": time.perf_counter() - job["t_submit"],
                    "harmony_iters": len(rec.get("clustering_iters", [])),
                    "clustering_iters": list(rec.get("clustering_iters", [])),
                    "footprint_bytes": job["need"], "n_cells": int(job["X"].shape[0])}
            fut.set_result((out, info))
            ok = True
        except Exception as e:  # noqa: BLE001 - reported through the future
            fut.set_exception(e)
            ok = False
        with self._cv:
            self._in_use -= job["need"]
            self._running -= 1
            self.stats["completed" if ok else "failed"] += 1
            self._cv.notify_all()


def latency_summary(latencies) -> dict:
    lat = np.sort(np.asarray(latencies, dtype=np.float64))
    if lat.size == 0:
        return {"p50_s": None, "p99_s": None, "n": 0}
    return {"p50_s": float(np.percentile(lat, 50)), "p99_s": float(np.percentile(lat, 99)),
            "mean_s": float(lat.mean()), "max_s": float(lat[-1]), "n": int(lat.size)}
