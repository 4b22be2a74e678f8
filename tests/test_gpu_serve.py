"""Served form: jobs that share the device give the same results as alone.

Several integration jobs of uneven size run through ``HarmonyServer`` with one
slot (one job at a time) and with four slots (all submitted at once, their
kernels overlapping on the device).  Every kernel of the path is
deterministic, so each job's Z_hat must be bitwise identical in both runs; one
job is also checked against the oracle.
"""

from __future__ import annotations

import numpy as np
import pytest

from oracle import harmony_oracle as ho
from rapids_singlecell_amd.datasets import synthetic_pcs
from rapids_singlecell_amd.serve import HarmonyServer, latency_summary

pytestmark = pytest.mark.gpu

SHAPES = [(3000, 20, 3, np.float32, "native"), (9000, 30, 5, np.float32, "split_bf16"),
          (1500, 16, 2, np.float64, "native"), (6000, 30, 4, np.float32, "native")]


def _jobs():
    jobs = []
    for i, (n, d, B, dt, prec) in enumerate(SHAPES):
        X, obs = synthetic_pcs(n, d, 6, B, seed=40 + i)
        codes, cats, _ = ho.batch_codes(obs, "batch")
        jobs.append((X.astype(dt), codes, len(cats),
                     dict(n_clusters=12, max_iter_harmony=3, max_iter_clustering=6,
                          random_state=i, precision=prec)))
    return jobs


def _run(slots, jobs):
    with HarmonyServer(slots=slots) as srv:
        futs = [srv.submit(X, c, B, **p) for X, c, B, p in jobs]
        res = [f.result(timeout=120) for f in futs]
    return res


def test_served_results_match_alone(device):
    jobs = _jobs()
    alone = _run(1, jobs)
    mixed = _run(4, jobs)
    for (za, ia), (zm, im) in zip(alone, mixed):
        assert za.shape == zm.shape and np.all(np.isfinite(za))
        np.testing.assert_array_equal(za, zm)
        assert ia["harmony_iters"] == im["harmony_iters"]
    lat = latency_summary([i["latency_s"] for _, i in mixed])
    assert lat["n"] == len(jobs) and lat["p99_s"] >= lat["p50_s"] > 0


def test_served_job_matches_oracle(device):
    X, obs = synthetic_pcs(2500, 16, 5, 3, seed=77)
    X = X.astype(np.float64)
    codes, cats, _ = ho.batch_codes(obs, "batch")
    Z = ho.l2_normalize(X)
    Y0 = Z[np.random.default_rng(1).choice(len(X), 10, replace=False)]
    kw = dict(n_clusters=10, max_iter_harmony=2, max_iter_clustering=5, random_state=3)
    with HarmonyServer(slots=2) as srv:
        z, info = srv.submit(X, codes, len(cats), init_centroids=Y0, **kw).result(timeout=120)
    zr = ho.harmonize(X, codes, len(cats), Y0=Y0, **kw)
    assert np.linalg.norm(z - zr) / np.linalg.norm(zr) < 1e-8


def test_served_rejects_oversized_job(device):
    with HarmonyServer(slots=1, memory_budget=1 << 20) as srv:
        X = np.zeros((50000, 50), np.float32)
        fut = srv.submit(X, np.zeros(50000, np.int32), 1)
        with pytest.raises(MemoryError):
            fut.result(timeout=10)
        assert srv.stats["rejected"] == 1
