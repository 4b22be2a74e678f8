"""Launch-shape autotuner (GPU).

``tune.autotune`` times the candidate sweep and centroid-pass workgroup counts
on a synthetic job of the requested shape and caches the fastest setting per
shape bucket, in memory and in a JSON file.  The block sweep's results do not
depend on its workgroup count (every block sum and objective term is fixed
point), so a Harmony run with only ``sweep_wg`` changed is bitwise the same;
the centroid-pass workgroup count changes the float summation order of the
centroid partials only.
"""

from __future__ import annotations

import json

import numpy as np
import pytest
import torch

from rapids_singlecell_amd import _native, tune
from rapids_singlecell_amd.datasets import synthetic_pcs
from rapids_singlecell_amd.preprocessing._harmony import harmonize

pytestmark = pytest.mark.gpu

SHAPE = (20_000, 20, 16, 4)  # cells, PCs, clusters, batches


def _run(X, obs, Y0):
    Z = harmonize(X, obs, "batch", n_clusters=Y0.shape[0], init_centroids=Y0, random_state=0,
                  max_iter_harmony=2, max_iter_clustering=5)
    return Z.cpu().numpy()


def test_autotune_measures_once_and_caches(device, tmp_path, monkeypatch):
    monkeypatch.setenv("RSC_AMD_TUNE_CACHE", str(tmp_path / "tune.json"))
    monkeypatch.setattr(tune, "_memo", {})
    before = _native.get_tuning()
    best = tune.autotune(*SHAPE, precision="native")
    print(f"\nautotune {SHAPE}: {best}", flush=True)
    cus = torch.cuda.get_device_properties(0).multi_processor_count
    assert {k: best[k] for k in ("sweep_wg", "centroid_wg")} in tune.candidates(cus)
    assert best["ms_per_clustering_iter"] > 0
    assert best["ms_per_clustering_iter"] <= best["default_ms"]
    assert _native.get_tuning() == before  # the knobs are restored after measuring
    key = tune.shape_key(*SHAPE, "native")
    assert json.loads((tmp_path / "tune.json").read_text())[key] == best
    # a second call measures nothing: from memory, then from the file
    monkeypatch.setattr(tune, "_time_setting",
                        lambda *a, **k: pytest.fail("a cached shape was measured again"))
    assert tune.autotune(*SHAPE, precision="native") == best
    monkeypatch.setattr(tune, "_memo", {})
    assert tune.autotune(*SHAPE, precision="native") == best


def test_tuned_settings_keep_results(device):
    n, d, K, B = SHAPE
    X, obs = synthetic_pcs(n, d, 6, B, seed=3)
    Xt = torch.as_tensor(X.astype(np.float32), device="cuda")
    Zn = X / np.linalg.norm(X, axis=1, keepdims=True)
    Y0 = Zn[np.random.default_rng(3).choice(n, K, replace=False)].astype(np.float32)
    before = _native.get_tuning()
    cus = torch.cuda.get_device_properties(0).multi_processor_count
    try:
        ref = _run(Xt, obs, Y0)
        for sw in sorted({cus, (3 * cus) // 4, cus // 2}):
            tune.apply({"sweep_wg": sw})
            np.testing.assert_array_equal(_run(Xt, obs, Y0), ref)
        tune.apply({"sweep_wg": cus // 2, "centroid_wg": 4 * cus})
        got = _run(Xt, obs, Y0)
        err = np.abs(got - ref).max() / np.abs(ref).max()
        print(f"\ncentroid_wg {4 * cus} vs automatic: max rel diff {err:.2e}", flush=True)
        assert err <= 1e-4
    finally:
        _native.set_tuning(**before)
