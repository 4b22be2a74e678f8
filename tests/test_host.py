"""Host-side logic of the drop-in, on the CPU: argument validation (same
exceptions as rapids_singlecell's ``harmonize`` / ``harmony_integrate``), batch
bookkeeping and the Harmony convergence rule.  No device work happens here: the
product path refuses CPU tensors instead of falling back to the CPU."""

from __future__ import annotations

import warnings

import numpy as np
import pandas as pd
import pytest
import torch

from oracle import harmony_oracle as ho
from rapids_singlecell_amd.datasets import SimpleAnnData
from rapids_singlecell_amd.preprocessing import harmony_integrate
from rapids_singlecell_amd.preprocessing._harmony import _is_convergent_harmony, harmonize
from rapids_singlecell_amd.preprocessing._harmony._helper import (
    _create_category_index_mapping,
    _get_batch_codes,
    _get_theta_array,
)


def _obs(n=60, B=3, seed=0):
    rng = np.random.default_rng(seed)
    return pd.DataFrame({"batch": rng.choice([f"b{i}" for i in range(B)], n),
                         "donor": rng.integers(0, 2, n)})


def _Z(n=60, d=5):
    return torch.randn(n, d, dtype=torch.float64)


# ---- harmonize: parameter errors (_harmony/__init__.py:155-200) -------------


@pytest.mark.parametrize("alpha", [0.0, -1.0, float("inf"), float("nan")])
def test_bad_alpha(alpha):
    with pytest.raises(ValueError, match="alpha must be a finite positive"):
        harmonize(_Z(), _obs(), "batch", alpha=alpha)


@pytest.mark.parametrize("thr", [-0.1, 1.5])
def test_bad_prune_threshold(thr):
    with pytest.raises(ValueError, match="batch_prune_threshold"):
        harmonize(_Z(), _obs(), "batch", batch_prune_threshold=thr)


def test_harmony1_ignores_alpha_checks():
    # dynamic_lambda=False skips the alpha check; the next stop is the device check
    with pytest.raises(TypeError, match="HIP device"):
        harmonize(_Z(), _obs(), "batch", alpha=-1.0, dynamic_lambda=False)


def test_bad_max_iter_and_method_and_colsum():
    with pytest.raises(ValueError, match="max_iter_harmony"):
        harmonize(_Z(), _obs(), "batch", max_iter_harmony=0)
    with pytest.raises(ValueError, match="correction_method"):
        harmonize(_Z(), _obs(), "batch", correction_method="exact")
    with pytest.raises(AssertionError):
        harmonize(_Z(), _obs(), "batch", colsum_algo="magic")
    with pytest.raises(AssertionError):
        harmonize(_Z(), _obs(), "batch", block_proportion=0.0)


def test_theta_size_mismatch():
    with pytest.raises(ValueError, match="must match number of batches"):
        harmonize(_Z(), _obs(B=3), "batch", theta=[1.0, 2.0])


def test_precision_checks():
    with pytest.raises(ValueError, match="precision"):
        harmonize(_Z(), _obs(), "batch", precision="fp8")
    with pytest.raises(ValueError, match="float32"):
        harmonize(_Z(), _obs(), "batch", precision="split_bf16")


def test_missing_batch_values():
    obs = _obs()
    obs.loc[3, "batch"] = None
    with pytest.raises(ValueError, match="missing values"):
        harmonize(_Z(), obs, "batch")


def test_no_cpu_fallback():
    with pytest.raises(TypeError, match="HIP device"):
        harmonize(_Z(), _obs(), "batch")
    with pytest.raises(TypeError):
        harmonize(np.zeros((60, 5)), _obs(), "batch")


# ---- harmony_integrate: wrapper checks (_harmony_integrate.py:155-250) -----


def test_integrate_bad_flavor_and_basis():
    ad = SimpleAnnData(obsm={"X_pca": np.zeros((60, 5))}, obs=_obs())
    with pytest.raises(ValueError, match="flavor"):
        harmony_integrate(ad, "batch", flavor="harmony3")
    with pytest.raises(ValueError, match="not available in adata.obsm"):
        harmony_integrate(ad, "batch", basis="X_umap")


def test_integrate_flavor_warnings():
    ad = SimpleAnnData(obsm={}, obs=_obs())
    with pytest.warns(UserWarning, match="ridge_lambda is ignored"):
        with pytest.raises(ValueError):
            harmony_integrate(ad, "batch", ridge_lambda=2.0)
    with warnings.catch_warnings(record=True) as w:
        warnings.simplefilter("always")
        with pytest.raises(ValueError):
            harmony_integrate(ad, "batch", flavor="harmony1", alpha=0.5,
                              batch_prune_threshold=0.1)
    msgs = [str(x.message) for x in w]
    assert any("alpha is ignored" in m for m in msgs)
    assert any("batch_prune_threshold is ignored" in m for m in msgs)


# ---- bookkeeping helpers agree with the oracle -----------------------------


@pytest.mark.parametrize("key", ["batch", ["batch", "donor"], ["donor"]])
def test_batch_codes_match_oracle(key):
    obs = _obs(200)
    codes = _get_batch_codes(obs, key)
    oc, ocats, onb = ho.batch_codes(obs, key)
    np.testing.assert_array_equal(codes.cat.codes.values, oc)
    assert list(codes.cat.categories) == ocats
    np.testing.assert_array_equal(codes.value_counts(sort=False).values, onb)


def test_theta_array_matches_oracle():
    for th in [2, 1.5, [1.0, 2.0, 3.0], np.array([0.5, 0.5, 0.5])]:
        np.testing.assert_array_equal(_get_theta_array(th, 3, np.float64),
                                      ho.theta_array(th, 3, np.float64))
    with pytest.raises(ValueError):
        _get_theta_array("2", 3, np.float64)


def test_theta_bool_is_an_int():
    """The reference tests ``isinstance(theta, float | int)`` (_helper.py:286),
    which a bool passes: True broadcasts as 1.0 and False as 0.0."""
    np.testing.assert_array_equal(_get_theta_array(True, 3, np.float32), np.ones(3, np.float32))
    np.testing.assert_array_equal(_get_theta_array(False, 2, np.float64), np.zeros(2))
    np.testing.assert_array_equal(ho.theta_array(True, 3, np.float32), np.ones(3, np.float32))


def test_category_index_mapping_matches_oracle():
    cats = np.random.default_rng(1).integers(0, 5, 1000).astype(np.int32)
    off, idx = _create_category_index_mapping(cats, 5)
    ooff, oidx = ho.category_index_mapping(cats, 5)
    np.testing.assert_array_equal(np.asarray(off), ooff)
    np.testing.assert_array_equal(np.asarray(idx), oidx)


# ---- Harmony convergence rule (_harmony/__init__.py:722-734) ---------------


def test_is_convergent_harmony():
    assert not _is_convergent_harmony([10.0], tol=1e-4)
    assert not _is_convergent_harmony([10.0, 9.0], tol=1e-4)
    assert _is_convergent_harmony([10.0, 9.9999], tol=1e-4)
    assert _is_convergent_harmony([10.0, 10.5], tol=1e-4)  # objective went up
    assert _is_convergent_harmony([-10.0, -10.0005], tol=1e-4)
    assert not _is_convergent_harmony([-10.0, -11.0], tol=1e-4)
