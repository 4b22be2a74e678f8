"""Per-entry-point parity of the HIP library against the oracle (GPU).

Tolerances follow the reference's kernel tests (tests/test_harmony_kernels.py):
float32 atol 1e-5 / rtol 1e-4 for elementwise kernels, float64 1e-10; integer
results are compared exactly.
"""

from __future__ import annotations

import ctypes

import numpy as np
import pytest
import torch

from oracle import harmony_oracle as ho
from rapids_singlecell_amd import _native

pytestmark = pytest.mark.gpu

DTYPES = [np.float32, np.float64]


def _t(x, dev):
    return torch.as_tensor(np.ascontiguousarray(x), device=dev)


def _p(t):
    return ctypes.c_void_p(t.data_ptr())


def _s(dev):
    return ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)


class Hold:
    """Uploads arrays and keeps the tensors alive until the test ends (a freed
    temporary would be handed to the next allocation while a kernel reads it)."""

    def __init__(self, dev):
        self.dev = dev
        self.keep = []

    def __call__(self, x):
        t = x if isinstance(x, torch.Tensor) else _t(x, self.dev)
        self.keep.append(t)
        return _p(t)


def _tol(dtype):
    return (1e-5, 1e-4) if dtype == np.float32 else (1e-10, 1e-9)


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("n_rows,n_cols", [(100, 50), (1, 20), (500, 3), (257, 100)])
def test_l2_normalize(device, dtype, n_rows, n_cols):
    rng = np.random.default_rng(42)
    src = rng.standard_normal((n_rows, n_cols)).astype(dtype)
    src[0] = 0  # zero row stays zero (kernels_normalize.cuh clamp)
    s = _t(src, device)
    d = torch.empty_like(s)
    lib = _native.load()
    _native.check(lib.hz_l2_normalize(src.itemsize, _p(s), _p(d), n_rows, n_cols, _s(device)), "l2")
    out = d.cpu().numpy()
    assert np.all(out[0] == 0)
    atol = 1e-6 if dtype == np.float32 else 1e-12
    np.testing.assert_allclose(out, ho.l2_normalize(src.astype(np.float64)), atol=atol, rtol=1e-5)


@pytest.mark.parametrize("n", [1, 2, 7, 1000, 65537, 200_000])
@pytest.mark.parametrize("seed", [0, 1000003, 0xFFFFFFFF])
def test_block_order_bit_exact(device, n, seed):
    lib = _native.load()
    nb = lib.hz_block_order_workspace_bytes(n)
    ws = torch.empty(nb, dtype=torch.uint8, device=device)
    out = torch.empty(n, dtype=torch.int32, device=device)
    _native.check(lib.hz_block_order(n, seed, _p(out), _p(ws), nb, _s(device)), "block_order")
    np.testing.assert_array_equal(out.cpu().numpy(), ho.block_order(n, seed))


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("n_batches,n_clusters", [(3, 20), (10, 100), (1, 5)])
@pytest.mark.parametrize("stabilized", [True, False])
def test_penalty(device, dtype, n_batches, n_clusters, stabilized):
    rng = np.random.default_rng(123)
    E = (rng.random((n_batches, n_clusters)) * 10).astype(dtype)
    O = (rng.random((n_batches, n_clusters)) * 10).astype(dtype)
    O[0, 0] = 0  # absent batch stays finite (test_penalty_absent_batch)
    theta = (rng.random(n_batches) * 2 + 0.5).astype(dtype)
    e, o, th = _t(E, device), _t(O, device), _t(theta, device)
    pen = torch.empty_like(e)
    lib = _native.load()
    _native.check(lib.hz_penalty(E.itemsize, _p(e), _p(o), _p(th), _p(pen), n_batches, n_clusters,
                                 int(stabilized), _s(device)), "penalty")
    atol, rtol = _tol(dtype)
    out = pen.cpu().numpy()
    assert np.all(np.isfinite(out))
    np.testing.assert_allclose(out, ho.penalty(E, O, theta, stabilized), atol=atol, rtol=rtol)


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("n_rows,n_cols,n_src", [(200, 50, 200), (50, 100, 50), (100, 30, 200)])
def test_fused_pen_norm(device, dtype, n_rows, n_cols, n_src):
    H = Hold(device)
    rng = np.random.default_rng(99)
    n_batches = 4
    sim = rng.random((n_src, n_cols)).astype(dtype)
    pen = (rng.random((n_batches, n_cols)) + 0.1).astype(dtype)
    cats = rng.integers(0, n_batches, size=n_rows).astype(np.int32)
    idx = rng.choice(n_src, size=n_rows, replace=False).astype(np.int32)
    term = float(dtype(-2) / dtype(0.1))
    out = torch.empty((n_rows, n_cols), dtype=torch.float32 if dtype == np.float32 else torch.float64,
                      device=device)
    lib = _native.load()
    _native.check(lib.hz_fused_pen_norm(sim.itemsize, H(sim), H(pen),
                                        H(cats), H(idx), _p(out), term,
                                        n_rows, n_cols, _s(device)), "fused_pen_norm")
    raw = np.exp(dtype(term) * (1 - sim[idx])) * pen[cats]
    expected = raw / raw.sum(axis=1, keepdims=True)
    atol, rtol = _tol(dtype)
    np.testing.assert_allclose(out.cpu().numpy(), expected, atol=atol, rtol=rtol)


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("dynamic", [True, False])
@pytest.mark.parametrize("threshold", [1e-5, None])
def test_lambda_kb(device, dtype, dynamic, threshold):
    H = Hold(device)
    rng = np.random.default_rng(7)
    B, K = 6, 40
    E = (rng.random((B, K)) * 20).astype(dtype)
    O = (rng.random((B, K)) * 20).astype(dtype)
    O[1, :5] = 0
    E[1, :3] = 0
    N_b = np.array([100, 0, 5, 50, 1, 30], dtype=dtype)
    O[4] = 0
    O[4, 7] = 1
    lib = _native.load()
    out = torch.empty((B, K), dtype=torch.float32 if dtype == np.float32 else torch.float64,
                      device=device)
    _native.check(lib.hz_lambda_kb(E.itemsize, H(E), H(O),
                                   H(N_b), B, K, int(dynamic), 0.2,
                                   int(threshold is not None), threshold or 0.0, 1.0, _p(out),
                                   _s(device)), "lambda_kb")
    exp = ho.compute_lambda_kb(E, O, N_b, alpha=0.2, threshold=threshold, ridge_lambda=1.0,
                               dynamic_lambda=dynamic)
    np.testing.assert_array_equal(out.cpu().numpy(), exp)


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("n_batches", [5, 50, 200])
def test_inv_mats(device, dtype, n_batches):
    H = Hold(device)
    n_clusters = 10
    rng = np.random.default_rng(42)
    O = (rng.random((n_batches, n_clusters)) * 100).astype(dtype)
    lam = (rng.random((n_batches, n_clusters)) * 5 + 0.1).astype(dtype)
    lam[0, 0] = dtype(1e30)
    inv = torch.empty((n_clusters, n_batches + 1, n_batches + 1),
                      dtype=torch.float32 if dtype == np.float32 else torch.float64, device=device)
    lib = _native.load()
    _native.check(lib.hz_inv_mats(O.itemsize, H(O), H(lam), _p(inv),
                                  n_batches, n_clusters, _s(device)), "inv_mats")
    exp = ho.inv_mats(O.astype(np.float64), lam.astype(np.float64))
    atol = 1e-6 if dtype == np.float32 else 1e-12
    np.testing.assert_allclose(inv.cpu().numpy(), exp, atol=atol, rtol=1e-5)


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("n_cells,n_clusters,n_batches", [(500, 20, 3), (1000, 100, 10), (200, 5, 2)])
@pytest.mark.parametrize("stabilized", [True, False])
def test_objective(device, dtype, n_cells, n_clusters, n_batches, stabilized):
    H = Hold(device)
    rng = np.random.default_rng(42)
    R = (rng.random((n_cells, n_clusters)) + 0.01).astype(dtype)
    sim = rng.random((n_cells, n_clusters)).astype(dtype)
    O = (rng.random((n_batches, n_clusters)) * 10 + 0.1).astype(dtype)
    E = (rng.random((n_batches, n_clusters)) * 10 + 0.1).astype(dtype)
    theta = (rng.random(n_batches) * 2).astype(dtype)
    lib = _native.load()
    nb = lib.hz_objective_workspace_bytes()
    ws = torch.empty(nb, dtype=torch.uint8, device=device)
    res = ctypes.c_double(0)
    _native.check(lib.hz_objective(R.itemsize, H(R), H(sim),
                                   H(O), H(E), H(theta),
                                   0.1, n_cells, n_clusters, n_batches, int(stabilized),
                                   ctypes.byref(res), _p(ws), nb, _s(device)), "objective")
    exp = ho.objective(R.astype(np.float64), sim.astype(np.float64), O.astype(np.float64),
                       E.astype(np.float64), theta.astype(np.float64), 0.1, stabilized)
    rtol = 1e-5 if dtype == np.float32 else 1e-10
    np.testing.assert_allclose(res.value, exp, rtol=rtol)


def _membership(device, n, bs, cats, seed, n_iter):
    lib = _native.load()
    B = int(cats.max()) + 1
    nb = lib.hz_block_membership_workspace_bytes(n, bs, B, n_iter)
    assert nb > 0
    ws = torch.empty(nb, dtype=torch.uint8, device=device)
    c = _t(cats.astype(np.int32), device)
    blk = torch.full((n_iter, n), -1, dtype=torch.int32, device=device)
    order = torch.full((n_iter, n), -1, dtype=torch.int32, device=device)
    _native.check(lib.hz_block_membership(n, bs, B, _p(c), seed & 0xFFFFFFFF, n_iter, _p(blk),
                                          _p(order), _p(ws), nb, _s(device)), "membership")
    return blk.cpu().numpy(), order.cpu().numpy()


def _check_membership(n, bs, cats, seed, blk, order):
    B = int(cats.max()) + 1
    for y in range(blk.shape[0]):
        perm = ho.block_order(n, (seed + y) & 0xFFFFFFFF)
        rank = np.empty(n, np.int64)
        rank[perm] = np.arange(n)
        # bit-exact block of every cell: rank in the key-sorted order // block size
        np.testing.assert_array_equal(blk[y], rank // bs, err_msg=f"iteration {y}")
        # the grouped order is a permutation sorted by (block, batch)
        o = order[y]
        np.testing.assert_array_equal(np.sort(o), np.arange(n))
        key = (rank[o] // bs) * B + cats[o]
        assert np.all(np.diff(key) >= 0)


@pytest.mark.parametrize("n,B", [(200_000, 8), (1_300_000, 16)])
@pytest.mark.parametrize("seed", [0, 1000003, 0xFFFFFFF8])
def test_block_membership_bit_exact(device, n, B, seed):
    """The block membership the clustering loop builds (boundary keys +
    counting sort, 16 iterations per set of launches, seeds wrapping at 2^32)
    equals the reference's key-sorted order cut into blocks
    (clustering.cu:292-319) bit for bit."""
    rng = np.random.default_rng(seed % 1000)
    cats = rng.integers(0, B, n).astype(np.int32)
    bs = int(n * 0.05)
    blk, order = _membership(device, n, bs, cats, seed, 16)
    _check_membership(n, bs, cats, seed, blk, order)


def test_block_membership_many_segments(device):
    """20 blocks x 100 batches = 2000 segments: the membership kernels still
    fit, but the placement's LDS staging does not, so the cells are placed
    with one store each; same membership."""
    n, B = 200_000, 100
    cats = np.random.default_rng(5).integers(0, B, n).astype(np.int32)
    bs = int(n * 0.05)
    blk, order = _membership(device, n, bs, cats, 4242, 3)
    _check_membership(n, bs, cats, 4242, blk, order)


@pytest.mark.parametrize("n", [9_000_000, 17_000_000])
def test_block_membership_large(device, n):
    """Past 2^23 cells the boundary buckets' histogram takes the whole 64 KiB
    of LDS (9M cells); past 2^24 (17M) the fine buckets are counted with
    global atomics instead.  Both give the reference's membership."""
    B = 6
    cats = np.random.default_rng(n % 1000).integers(0, B, n).astype(np.int32)
    bs = int(n * 0.05)
    print(f"\nmembership n={n}", flush=True)
    blk, order = _membership(device, n, bs, cats, 12345, 2)
    _check_membership(n, bs, cats, 12345, blk, order)


def test_block_membership_key_order_fallback(device):
    """100 blocks x 41 batches = 4100 segments do not fit the membership
    kernels' LDS: the full key ranking is used instead; same membership."""
    n, B = 5000, 41
    cats = (np.arange(n) % B).astype(np.int32)
    blk, order = _membership(device, n, 50, cats, 77, 3)
    _check_membership(n, 50, cats, 77, blk, order)
