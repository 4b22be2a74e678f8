"""Block sweep: concurrent jobs and the bounded-wait error path (GPU).

The block sweep is one cooperative launch whose workgroups wait for each
other's block sums.  Two sweeps must never share the device's CUs, so the
library orders every sweep on a device after the previous one, whichever
thread and stream launched it.  The first test runs two sweep-sized jobs from
two threads on two streams at once and checks each against its result alone
(bitwise: every kernel of the path is deterministic).

The second forces the bounded wait to give up from the host (a spin limit of
one poll), so that every kernel still exits normally, and checks that the
call reports HZ_ERR_LAUNCH and that the next call on the same thread works.
"""

from __future__ import annotations

import threading

import numpy as np
import pytest
import torch

from oracle import harmony_oracle as ho
from rapids_singlecell_amd import _native
from rapids_singlecell_amd.datasets import synthetic_pcs
from rapids_singlecell_amd.preprocessing._harmony import harmonize

pytestmark = pytest.mark.gpu

FIXED = dict(max_iter_harmony=2, max_iter_clustering=6, tol_harmony=float("-inf"),
             tol_clustering=float("-inf"))


def _job(n, d, B, K, seed, precision):
    X, obs = synthetic_pcs(n, d, 8, B, seed=seed)
    Z = ho.l2_normalize(X.astype(np.float64))
    Y0 = Z[np.random.default_rng(seed).choice(n, K, replace=False)].astype(np.float32)
    return X, obs, dict(n_clusters=K, init_centroids=Y0, precision=precision,
                        random_state=seed, **FIXED)


def _run(X, obs, kw, stream=None):
    dev = torch.device("cuda", 0)
    if stream is None:
        out = harmonize(torch.as_tensor(X, device=dev), obs, "batch", **kw)
        return out.cpu().numpy()
    with torch.cuda.stream(stream):
        out = harmonize(torch.as_tensor(X, device=dev), obs, "batch", **kw)
        res = out.cpu().numpy()
    stream.synchronize()
    return res


def test_concurrent_sweeps_match_alone(device):
    jobs = [_job(120_000, 50, 8, 100, 11, "native"), _job(90_000, 30, 5, 60, 12, "split_bf16")]
    alone = [_run(*j) for j in jobs]
    results = [None, None]
    errors = []

    def worker(i):
        try:
            s = torch.cuda.Stream(device=device)
            for _ in range(2):  # overlap each job's sweeps with the other's
                results[i] = _run(*jobs[i], stream=s)
        except Exception as e:  # noqa: BLE001 - re-raised below
            errors.append(e)
        finally:
            _native.release_thread()

    ts = [threading.Thread(target=worker, args=(i,)) for i in range(2)]
    for t in ts:
        t.start()
    for t in ts:
        t.join(timeout=300)
    assert not errors, errors
    for a, r in zip(alone, results):
        np.testing.assert_array_equal(a, r)


def test_sweep_wait_limit_reports_launch_error(device):
    X, obs, kw = _job(40_000, 30, 4, 40, 21, "native")
    _native.set_tuning(spin_limit=1)
    try:
        with pytest.raises(_native.HarmonyNativeError, match="timed out"):
            _run(X, obs, kw)
    finally:
        _native.set_tuning(spin_limit=0)
    # the error was reported, every kernel exited, and the next call works
    got = _run(X, obs, kw)
    codes, cats, _ = ho.batch_codes(obs, "batch")
    Zr = ho.harmonize(X.astype(np.float64), codes, len(cats),
                      Y0=kw["init_centroids"].astype(np.float64), n_clusters=kw["n_clusters"],
                      random_state=kw["random_state"], **FIXED)
    err = np.linalg.norm(got - Zr) / np.linalg.norm(Zr)
    print(f"\nafter the forced wait error: rel err vs FP64 oracle = {err:.3e}", flush=True)
    assert err < 1e-4
