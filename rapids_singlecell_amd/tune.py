"""Autotuner of the library's launch shapes.

For one problem shape (cells, PCs, clusters, batches, precision) the tuner
builds a synthetic job of that shape on the device, and for every candidate
setting of the launch knobs

  * ``sweep_wg``    -- workgroups of the block sweep (one per CU at most),
  * ``centroid_wg`` -- workgroups of the centroid pass,

times two clustering iterations with the library's HIP-event profiler (after
one untimed iteration) and keeps the fastest.  Results are cached per shape
bucket in memory and in a JSON file (``RSC_AMD_TUNE_CACHE``, default
``~/.cache/rapids_singlecell_amd/tune.json``), so a shape is measured once per
machine.  The reference's analogue is the column-sum algorithm benchmark of
its Harmony helper (``_harmony/_helper.py:401-504``), which times candidate
kernels on the live data and caches the winner.

The knobs are process-wide library settings: ``apply`` sets them, and the
workspace of a job must be sized and used under the same settings.
"""

from __future__ import annotations

import json
import os
import threading
from pathlib import Path

import numpy as np
import torch

from . import _native

_lock = threading.Lock()
_memo: dict = {}


def _cache_path() -> Path:
    p = os.environ.get("RSC_AMD_TUNE_CACHE")
    if p:
        return Path(p)
    return Path.home() / ".cache" / "rapids_singlecell_amd" / "tune.json"


def shape_key(n_cells: int, n_pcs: int, n_clusters: int, n_batches: int, precision: str) -> str:
    """Shape bucket: cells rounded to powers of two (the launch shapes depend
    on the chunk counts, which scale with the cells), the rest exact."""
    nb = 1 << max(0, int(round(np.log2(max(n_cells, 1)))))
    dev = torch.cuda.get_device_name(0) if torch.cuda.is_available() else "cpu"
    return f"{dev}|n~{nb}|d{n_pcs}|K{n_clusters}|B{n_batches}|{precision}"


def _load_file() -> dict:
    try:
        return json.loads(_cache_path().read_text())
    except Exception:  # noqa: BLE001 - a missing or unreadable cache is empty
        return {}


def _save_file(entries: dict) -> None:
    path = _cache_path()
    try:
        path.parent.mkdir(parents=True, exist_ok=True)
        tmp = path.with_suffix(".tmp")
        tmp.write_text(json.dumps(entries, indent=1, sort_keys=True))
        tmp.replace(path)
    except OSError:
        pass  # a read-only home only loses the cache


def candidates(cus: int) -> list[dict]:
    out = []
    for sw in sorted({cus, (3 * cus) // 4, cus // 2}):
        for cw in (cus, 2 * cus, 4 * cus):
            out.append({"sweep_wg": sw, "centroid_wg": cw})
    return out


def _time_setting(X, codes, B, K, prec: int, setting: dict, iters: int = 2) -> float:
    from .preprocessing._harmony._job import HarmonyJob

    _native.set_tuning(**setting)
    N_b = np.bincount(codes, minlength=B).astype(np.float32)
    job = HarmonyJob(X, codes, B, n_clusters=K, theta_np=np.full(B, 2.0, np.float32),
                     N_b_np=N_b, max_iter_clustering=1, tol_clustering=float("-inf"),
                     stabilized_penalty=False, dynamic_lambda=False, precision=prec)
    job.init()
    job.clustering(0)  # untimed: first launches, block orders
    torch.cuda.synchronize()
    _native.profile_collect()
    _native.profile_enable(True)
    try:
        for i in range(iters):
            job.clustering(1 + i)
        prof = _native.profile_collect()
    finally:
        _native.profile_enable(False)
    del job
    kinds = ("sweep", "centroid", "yreduce", "row")
    return sum(prof[k][0] for k in kinds) / iters


def autotune(n_cells: int, n_pcs: int, n_clusters: int, n_batches: int,
             precision: str = "native", *, refresh: bool = False, verbose: bool = False) -> dict:
    """Best launch knobs for a shape (measured once, then cached)."""
    key = shape_key(n_cells, n_pcs, n_clusters, n_batches, precision)
    with _lock:
        if not refresh:
            if key in _memo:
                return dict(_memo[key])
            disk = _load_file()
            if key in disk:
                _memo[key] = disk[key]
                return dict(disk[key])
        lib = _native.load()
        before = _native.get_tuning()
        rng = np.random.default_rng(0)
        X = torch.as_tensor(rng.standard_normal((n_cells, n_pcs)).astype(np.float32),
                            device="cuda")
        codes = rng.integers(0, n_batches, n_cells).astype(np.int32)
        prec = {"native": 0, "split_bf16": 1}[precision]
        cus = torch.cuda.get_device_properties(0).multi_processor_count
        timings = []
        try:
            for setting in candidates(cus):
                try:
                    ms = _time_setting(X, codes, n_batches, n_clusters, prec, setting)
                except Exception as e:  # noqa: BLE001 - an unusable setting is skipped
                    if verbose:
                        print(f"[tune] {setting}: {e}")
                    continue
                timings.append((ms, setting))
                if verbose:
                    print(f"[tune] {setting}: {ms:.4f} ms per clustering iteration")
        finally:
            _native.set_tuning(**before)
        del lib
        if not timings:
            raise RuntimeError(f"no launch setting ran for {key}")
        ms, best = min(timings, key=lambda t: t[0])
        entry = dict(best, ms_per_clustering_iter=round(ms, 5),
                     default_ms=round(next((t for t, s in timings
                                            if s == {"sweep_wg": cus, "centroid_wg": 2 * cus}),
                                           float("nan")), 5))
        _memo[key] = entry
        disk = _load_file()
        disk[key] = entry
        _save_file(disk)
        return dict(entry)


def apply(setting: dict) -> None:
    """Set the library's launch knobs to a tuned setting."""
    _native.set_tuning(**{k: int(v) for k, v in setting.items() if k in _native.TUNE_KEYS})
