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
