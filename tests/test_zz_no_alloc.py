"""hz_block_order_noalloc: the block order without a workspace or any
allocation, against hz_block_order with a caller workspace (GPU).

Bit-exact (integer work).  Two calls in a row, the second on a smaller input,
so that what the first call left in the shared static scratch cannot leak into
the second result; the key order is also checked against the committed golden
vectors of the oracle's block_order.
"""

from __future__ import annotations

from pathlib import Path

import numpy as np
import pytest
import torch

from rapids_singlecell_amd import _native

pytestmark = pytest.mark.gpu

GOLDEN = Path(__file__).resolve().parent / "golden" / "block_order.npz"


def _with_workspace(lib, n, seed, stream):
    out = torch.empty(n, dtype=torch.int32, device="cuda")
    ws = torch.empty(lib.hz_block_order_workspace_bytes(n), dtype=torch.uint8, device="cuda")
    _native.check(lib.hz_block_order(n, seed, out.data_ptr(), ws.data_ptr(), ws.numel(), stream),
                  "hz_block_order")
    return out


def _no_alloc(lib, n, seed, stream):
    out = torch.empty(n, dtype=torch.int32, device="cuda")
    _native.check(lib.hz_block_order_noalloc(n, seed, out.data_ptr(), stream),
                  "hz_block_order_noalloc")
    return out


def test_no_alloc_matches_workspace_entry_point(device):
    lib = _native.load()
    stream = torch.cuda.current_stream().cuda_stream
    # large first (fills most of the scratch), then smaller, then a size
    # whose bucket count differs again; one seed wraps uint32
    for n, seed in [(1_300_000, 0xFFFFFFF0), (200_000, 1000003), (4097, 42)]:
        print(f"\nno-alloc block order n={n} seed={seed}", flush=True)
        ref = _with_workspace(lib, n, seed, stream)
        got = _no_alloc(lib, n, seed, stream)
        torch.cuda.synchronize()
        assert torch.equal(ref, got), (n, seed)
        # a permutation of 0..n-1
        assert torch.equal(torch.sort(got.long()).values,
                           torch.arange(n, device=got.device))


def test_no_alloc_golden_vectors(device):
    lib = _native.load()
    stream = torch.cuda.current_stream().cuda_stream
    with np.load(GOLDEN) as g:
        assert g.files
        for key in g.files:  # keys "n<cells>_s<seed>"
            n = int(key.split("_")[0][1:])
            seed = int(key.split("_")[1][1:])
            got = _no_alloc(lib, n, seed, stream).cpu().numpy()
            np.testing.assert_array_equal(got, g[key])


def test_no_alloc_rejects_oversized_input(device):
    lib = _native.load()
    out = torch.empty(1, dtype=torch.int32, device="cuda")
    n = 8_000_000
    assert lib.hz_block_order_workspace_bytes(n) > (32 << 20)
    rc = lib.hz_block_order_noalloc(n, 0, out.data_ptr(), None)
    assert rc == -5  # HZ_ERR_WORKSPACE, returned before any launch
    assert b"static scratch" in lib.hz_last_error()
