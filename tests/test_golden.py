"""Golden fixtures (tests/golden/*.npz, written by tests/golden/make_golden.py).

CPU: the oracle still reproduces every fixture (guards the oracle against
drift).  GPU: the HIP path reproduces them through the C-ABI / drop-in --
block orders bit-exactly, lambda_kb exactly, FP64 Harmony within 1e-8 relative
with identical clustering iteration counts.
"""

from __future__ import annotations

import ctypes
from pathlib import Path

import numpy as np
import pandas as pd
import pytest

from oracle import harmony_oracle as ho

GOLDEN = Path(__file__).resolve().parent / "golden"


def _load(name):
    return np.load(GOLDEN / name, allow_pickle=False)


def _order_cases():
    with _load("block_order.npz") as f:
        return [(k, int(k.split("_")[0][1:]), int(k.split("_")[1][1:])) for k in f.files]


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


HARMONY_KW = dict(n_clusters=8, max_iter_harmony=3, max_iter_clustering=10,
                  tol_harmony=float("-inf"), tol_clustering=1e-5, random_state=7)


# ---------------------------------------------------------------- CPU


@pytest.mark.parametrize("key,n,seed", _order_cases())
def test_oracle_block_order_fixture(key, n, seed):
    with _load("block_order.npz") as f:
        np.testing.assert_array_equal(ho.block_order(n, seed), f[key])


def test_oracle_lambda_fixture():
    with _load("lambda_kat.npz") as f:
        np.testing.assert_array_equal(
            ho.compute_lambda_kb(f["prune_E"], f["prune_O"], f["prune_Nb"], alpha=0.2,
                                 threshold=1e-5, ridge_lambda=1.0, dynamic_lambda=True),
            f["prune_out"])


@pytest.mark.parametrize("flavor", ["harmony1", "harmony2"])
def test_oracle_harmony_fixture(flavor):
    with _load("harmony_small.npz") as f:
        stab = flavor == "harmony2"
        rec = {}
        Z = ho.harmonize(f["X"], f["cats"], int(f["n_batches"]), Y0=f["Y0"],
                         stabilized_penalty=stab, dynamic_lambda=stab, record=rec, **HARMONY_KW)
        assert rec["clustering_iters"] == list(f[f"{flavor}_clustering_iters"])
        assert _rel(Z, f[f"{flavor}_Z_hat"]) < 1e-12


# ---------------------------------------------------------------- GPU


@pytest.mark.gpu
@pytest.mark.parametrize("key,n,seed", _order_cases())
def test_gpu_block_order_fixture(device, key, n, seed):
    import torch

    from rapids_singlecell_amd import _native

    lib = _native.load()
    nb = lib.hz_block_order_workspace_bytes(n)
    ws = torch.empty(nb, dtype=torch.uint8, device=device)
    out = torch.empty(n, dtype=torch.int32, device=device)
    s = ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)
    _native.check(lib.hz_block_order(n, seed, ctypes.c_void_p(out.data_ptr()),
                                     ctypes.c_void_p(ws.data_ptr()), nb, s), "block_order")
    with _load("block_order.npz") as f:
        np.testing.assert_array_equal(out.cpu().numpy(), f[key])


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["prune", "ridge0", "zden"])
def test_gpu_lambda_fixture(device, case):
    import torch

    from rapids_singlecell_amd import _native

    args = {"prune": (1, 0.2, 1, 1e-5, 1.0), "ridge0": (0, 0.2, 0, 0.0, 0.0),
            "zden": (1, 0.2, 0, 0.0, 1.0)}[case]
    with _load("lambda_kat.npz") as f:
        E, O, Nb, want = (torch.as_tensor(f[f"{case}_{k}"], device=device)
                          for k in ("E", "O", "Nb", "out"))
    B, K = E.shape
    out = torch.empty_like(E)
    lib = _native.load()
    p = lambda t: ctypes.c_void_p(t.data_ptr())  # noqa: E731
    _native.check(lib.hz_lambda_kb(8, p(E), p(O), p(Nb), B, K, args[0], args[1], args[2],
                                   args[3], args[4], p(out),
                                   ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)),
                  "lambda_kb")
    np.testing.assert_array_equal(out.cpu().numpy(), want.cpu().numpy())


@pytest.mark.gpu
@pytest.mark.parametrize("flavor", ["harmony1", "harmony2"])
def test_gpu_harmony_fixture(device, flavor):
    import torch

    from rapids_singlecell_amd.preprocessing._harmony import harmonize

    with _load("harmony_small.npz") as f:
        stab = flavor == "harmony2"
        obs = pd.DataFrame({"batch": pd.Categorical(f["cats"])})
        rec = {}
        Z = harmonize(torch.as_tensor(f["X"], device=device), obs, "batch",
                      init_centroids=f["Y0"], stabilized_penalty=stab, dynamic_lambda=stab,
                      record=rec, **HARMONY_KW)
        assert rec["clustering_iters"] == list(f[f"{flavor}_clustering_iters"])
        np.testing.assert_allclose(rec["objectives"], f[f"{flavor}_objectives"], rtol=1e-9)
        assert _rel(Z.cpu().numpy(), f[f"{flavor}_Z_hat"]) < 1e-8
