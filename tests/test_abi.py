"""C-ABI checks that need no GPU: the library loads, exports every entry point
that ``include/harmony_hip.h`` declares, and its host-only queries (version,
workspace sizes, tuning knobs, argument validation) behave."""

from __future__ import annotations

import ctypes
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "harmony_hip.h"


def _declared() -> list[str]:
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^[A-Za-z_][\w\s\*]*?\b(hz_\w+)\s*\(", text, flags=re.M)))


@pytest.fixture(scope="module")
def lib():
    from rapids_singlecell_amd import _build, _native

    if not _native.lib_path().exists():
        _build.build()
    return _native.load()


def test_header_declares_entry_points():
    names = _declared()
    assert "hz_clustering_loop" in names and "hz_correction" in names
    assert len(names) >= 20


def test_library_exports_every_declared_symbol(lib):
    missing = [n for n in _declared() if not hasattr(lib, n)]
    assert not missing, f"declared in {HEADER.name} but not exported: {missing}"


def test_loader_signatures_cover_header():
    from rapids_singlecell_amd import _native

    assert set(_declared()) <= set(_native.exported_symbols())


def test_version_and_error_string(lib):
    assert lib.hz_version() >= 1
    # a rejected argument sets the thread's last error without touching a device
    rc = lib.hz_l2_normalize(4, None, None, -1, 3, None)
    assert rc == -1
    assert b"bad shape" in lib.hz_last_error()


def _dims(n, d, K, B, bs, esz=4, prec=0):
    from rapids_singlecell_amd import _native

    return _native.dims(n, d, K, B, bs, esz, prec)


def test_workspace_bytes_host_only(lib):
    small = lib.hz_workspace_bytes(ctypes.byref(_dims(2000, 20, 10, 3, 100)))
    big = lib.hz_workspace_bytes(ctypes.byref(_dims(200_000, 50, 100, 8, 10_000)))
    assert 0 < small < big
    assert small % 256 == 0 and big % 256 == 0
    # unsupported / invalid shapes report 0 and a reason
    assert lib.hz_workspace_bytes(ctypes.byref(_dims(0, 20, 10, 3, 1))) == 0
    assert lib.hz_workspace_bytes(ctypes.byref(_dims(100, 20, 10, 3, 1, esz=2))) == 0
    assert lib.hz_last_error()
    assert lib.hz_block_order_workspace_bytes(0) == 0
    assert lib.hz_block_order_workspace_bytes(1000) > 0
    assert lib.hz_objective_workspace_bytes() > 0
    assert lib.hz_kmeans_workspace_bytes(4, 1000, 20, 10) > 0


def test_tuning_roundtrip(lib):
    from rapids_singlecell_amd import _native

    before = _native.get_tuning()
    try:
        _native.set_tuning(row_chunk=32)
        assert _native.get_tuning()["row_chunk"] == 32
        with pytest.raises(Exception):
            _native.set_tuning(row_chunk=17)  # must be a multiple of 16
    finally:
        _native.set_tuning(**before)
    assert _native.get_tuning() == before


def test_batched_entry_point_validates_on_the_host(lib):
    """hz_clustering_loop_batched refuses an empty batch, unequal max_iter and
    jobs of different kernel shapes before any launch (the workspaces are
    only carved, never touched, so aligned placeholder addresses suffice)."""
    from rapids_singlecell_amd.preprocessing._harmony._batch import HzJob

    assert lib.hz_batched_workspace_bytes(0, 10) == 0
    assert lib.hz_batched_workspace_bytes(4, 20) > lib.hz_batched_workspace_bytes(2, 20)
    assert lib.hz_clustering_loop_batched(None, 0, None, 0, None) == -1
    arr = (HzJob * 2)()
    for i, d in enumerate((30, 100)):
        dm = _dims(5000, d, 40, 3, 250)
        arr[i].dims = dm
        arr[i].max_iter = 4
        arr[i].ws = 1 << 40
        arr[i].ws_bytes = lib.hz_workspace_bytes(ctypes.byref(dm))
        arr[i].Z_norm = 1 << 41
        arr[i].R = 1 << 42
    p = ctypes.cast(arr, ctypes.c_void_p)
    arr[1].max_iter = 5
    assert lib.hz_clustering_loop_batched(p, 2, None, 0, None) == -1
    assert b"same max_iter" in lib.hz_last_error()
    arr[1].max_iter = 4
    assert lib.hz_clustering_loop_batched(p, 2, None, 0, None) == -4
    assert b"kernel shape" in lib.hz_last_error()
