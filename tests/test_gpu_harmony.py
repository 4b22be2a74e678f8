"""Job-level parity of the HIP Harmony path against the oracle (GPU).

The reference's clustering is a chaotic fixed-point iteration, so runs compare
at fixed iteration counts (``tol = -inf``) as well as with the reference's own
convergence rules.  Initial centroids are injected (the reference's cuML
k-means is not reproducible here).
"""

from __future__ import annotations

import ctypes

import numpy as np
import pytest
import torch

from oracle import harmony_oracle as ho
from rapids_singlecell_amd import _native
from rapids_singlecell_amd.datasets import SimpleAnnData, synthetic_pcs
from rapids_singlecell_amd.preprocessing._harmony import harmonize

pytestmark = pytest.mark.gpu

_U32 = 0xFFFFFFFF


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


def _tdt(dtype):
    return torch.float32 if dtype == np.float32 else torch.float64


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - np.asarray(b, np.float64))
                 / np.linalg.norm(np.asarray(b, np.float64)))


class Problem:
    """Inputs shared by the oracle and the GPU for one synthetic case."""

    def __init__(self, n=2000, d=20, n_types=6, B=3, K=12, seed=3, dtype=np.float64,
                 dirichlet=False):
        X, obs = synthetic_pcs(n, d, n_types, B, seed=seed, dirichlet_batches=dirichlet)
        self.obs = obs
        self.dtype = dtype
        self.X = X.astype(dtype)
        self.codes, cats, self.n_b = ho.batch_codes(obs, "batch")
        self.B = len(cats)
        self.n, self.d, self.K = n, d, K
        self.Z_norm = ho.l2_normalize(self.X)
        rng = np.random.default_rng(seed + 100)
        self.Y0 = self.Z_norm[rng.choice(n, K, replace=False)].copy()
        N_b = self.n_b.astype(dtype)
        self.N_b = N_b
        self.Pr_b = (N_b.reshape(-1, 1) / n).astype(dtype)
        self.theta = ho.theta_array(2.0, self.B, dtype)
        self.offsets, self.cell_idx = ho.category_index_mapping(self.codes, self.B)

    def dims(self, block_proportion=0.05):
        return _native.dims(self.n, self.d, self.K, self.B, int(self.n * block_proportion),
                            np.dtype(self.dtype).itemsize)


def _gpu_init(pb, dev, stabilized):
    lib = _native.load()
    dm = pb.dims()
    nb = lib.hz_workspace_bytes(ctypes.byref(dm))
    assert nb > 0, lib.hz_last_error()
    ws = torch.empty(nb, dtype=torch.uint8, device=dev)
    tdt = _tdt(pb.dtype)
    R = torch.empty((pb.n, pb.K), dtype=tdt, device=dev)
    E = torch.empty((pb.B, pb.K), dtype=tdt, device=dev)
    O = torch.empty((pb.B, pb.K), dtype=tdt, device=dev)
    obj = torch.empty(1, dtype=tdt, device=dev)
    bufs = dict(Z=_t(pb.Z_norm, dev), Y0=_t(pb.Y0, dev), cats=_t(pb.codes, dev),
                off=_t(pb.offsets, dev), ci=_t(pb.cell_idx, dev),
                Pr=_t(pb.Pr_b.ravel(), dev), th=_t(pb.theta, dev), Nb=_t(pb.N_b, dev))
    _native.check(lib.hz_init_clusters(ctypes.byref(dm), _p(bufs["Z"]), _p(bufs["Y0"]),
                                       _p(bufs["cats"]), _p(bufs["off"]), _p(bufs["ci"]),
                                       _p(bufs["Pr"]), _p(bufs["th"]), 0.1, int(stabilized),
                                       _p(R), _p(E), _p(O), _p(obj), _p(ws), nb, _s(dev)), "init")
    return lib, dm, ws, nb, bufs, R, E, O, obj


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("stabilized", [True, False])
def test_init_clusters(device, dtype, stabilized):
    pb = Problem(dtype=dtype)
    _, _, _, _, _, R, E, O, obj = _gpu_init(pb, device, stabilized)
    Rr, Er, Or, objr = ho.init_clusters(pb.Z_norm.astype(np.float64), pb.Y0.astype(np.float64),
                                       sigma=0.1, Pr_b=pb.Pr_b.astype(np.float64),
                                       theta=pb.theta.astype(np.float64), cats=pb.codes,
                                       n_batches=pb.B, stabilized=stabilized)
    tol = 2e-4 if dtype == np.float32 else 1e-10
    np.testing.assert_allclose(R.cpu().numpy(), Rr, atol=tol * 1e-2, rtol=tol)
    np.testing.assert_allclose(O.cpu().numpy(), Or, rtol=tol, atol=tol)
    np.testing.assert_allclose(E.cpu().numpy(), Er, rtol=tol, atol=tol)
    np.testing.assert_allclose(float(obj.item()), objr, rtol=tol)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("stabilized", [True, False])
@pytest.mark.parametrize("n,block_proportion", [(2000, 0.05), (1999, 0.05), (3001, 0.3)])
def test_clustering_fixed_iters(device, dtype, stabilized, n, block_proportion):
    """Three clustering iterations without early stop; R/E/O/objective vs oracle."""
    pb = Problem(n=n, dtype=dtype)
    lib, _, ws, nb, bufs, R, E, O, obj = _gpu_init(pb, device, stabilized)
    dm = pb.dims(block_proportion)
    nb = lib.hz_workspace_bytes(ctypes.byref(dm))
    ws = torch.empty(nb, dtype=torch.uint8, device=device)
    Rr, Er, Or, _ = ho.init_clusters(pb.Z_norm, pb.Y0, sigma=0.1, Pr_b=pb.Pr_b, theta=pb.theta,
                                     cats=pb.codes, n_batches=pb.B, stabilized=stabilized)
    # start both from the oracle's state so the comparison isolates the loop
    R.copy_(_t(Rr, device))
    E.copy_(_t(Er, device))
    O.copy_(_t(Or, device))
    seed = 12345
    last = torch.zeros(1, dtype=_tdt(dtype), device=device)
    n_iter = ctypes.c_int(0)
    _native.check(lib.hz_clustering_loop(ctypes.byref(dm), _p(bufs["Z"]), _p(R), _p(E), _p(O),
                                         _p(bufs["Pr"]), _p(bufs["cats"]), _p(bufs["th"]), 0.1,
                                         float("-inf"), 3, seed, int(stabilized), _p(last),
                                         ctypes.byref(n_iter), _p(ws), nb, _s(device)), "loop")
    last_r, n_r, objs = ho.clustering_loop(pb.Z_norm, Rr, Er, Or, pb.Pr_b, pb.codes, pb.theta,
                                           sigma=0.1, tol=float("-inf"), max_iter=3,
                                           block_size=int(n * block_proportion), seed=seed,
                                           stabilized=stabilized)
    assert n_iter.value == n_r == 3
    tol = 1e-3 if dtype == np.float32 else 1e-8
    assert rel(R.cpu().numpy(), Rr) < tol
    assert rel(O.cpu().numpy(), Or) < tol
    assert rel(E.cpu().numpy(), Er) < tol
    np.testing.assert_allclose(float(last.item()), last_r, rtol=tol)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_clustering_past_one_order_batch(device, dtype):
    """20 clustering iterations: the block orders of iterations 0-15 and 16-19
    come from two separate batched builds.  FP64 must match the oracle (same
    permutation every iteration); FP32 (block sweep) keeps the bookkeeping
    invariants: rows of R sum to 1 and sum_b O = column sums of R."""
    pb = Problem(n=2000, dtype=dtype)
    lib, dm, ws, nb, bufs, R, E, O, obj = _gpu_init(pb, device, False)
    Rr, Er, Or, _ = ho.init_clusters(pb.Z_norm, pb.Y0, sigma=0.1, Pr_b=pb.Pr_b, theta=pb.theta,
                                     cats=pb.codes, n_batches=pb.B, stabilized=False)
    R.copy_(_t(Rr, device))
    E.copy_(_t(Er, device))
    O.copy_(_t(Or, device))
    last = torch.zeros(1, dtype=_tdt(dtype), device=device)
    n_iter = ctypes.c_int(0)
    _native.check(lib.hz_clustering_loop(ctypes.byref(dm), _p(bufs["Z"]), _p(R), _p(E), _p(O),
                                         _p(bufs["Pr"]), _p(bufs["cats"]), _p(bufs["th"]), 0.1,
                                         float("-inf"), 20, 99, 0, _p(last),
                                         ctypes.byref(n_iter), _p(ws), nb, _s(device)), "loop")
    assert n_iter.value == 20
    Rg, Og = R.cpu().numpy().astype(np.float64), O.cpu().numpy().astype(np.float64)
    if dtype == np.float64:
        ho.clustering_loop(pb.Z_norm, Rr, Er, Or, pb.Pr_b, pb.codes, pb.theta, sigma=0.1,
                           tol=float("-inf"), max_iter=20, block_size=int(pb.n * 0.05), seed=99,
                           stabilized=False)
        assert rel(Rg, Rr) < 1e-7
        assert rel(Og, Or) < 1e-7
    else:
        np.testing.assert_allclose(Rg.sum(axis=1), 1.0, atol=1e-5)
        np.testing.assert_allclose(Og.sum(axis=0), Rg.sum(axis=0), rtol=1e-4)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_clustering_convergence_rule(device, dtype):
    """Window-3 convergence (clustering.cu:380-387): same iteration count."""
    pb = Problem(n=3000, dtype=dtype, K=10)
    lib, dm, ws, nb, bufs, R, E, O, obj = _gpu_init(pb, device, True)
    Rr, Er, Or, _ = ho.init_clusters(pb.Z_norm, pb.Y0, sigma=0.1, Pr_b=pb.Pr_b, theta=pb.theta,
                                     cats=pb.codes, n_batches=pb.B, stabilized=True)
    R.copy_(_t(Rr, device))
    E.copy_(_t(Er, device))
    O.copy_(_t(Or, device))
    last = torch.zeros(1, dtype=_tdt(dtype), device=device)
    n_iter = ctypes.c_int(0)
    tol = 1e-3
    _native.check(lib.hz_clustering_loop(ctypes.byref(dm), _p(bufs["Z"]), _p(R), _p(E), _p(O),
                                         _p(bufs["Pr"]), _p(bufs["cats"]), _p(bufs["th"]), 0.1,
                                         tol, 50, 7, 1, _p(last), ctypes.byref(n_iter), _p(ws),
                                         nb, _s(device)), "loop")
    _, n_r, _ = ho.clustering_loop(pb.Z_norm, Rr, Er, Or, pb.Pr_b, pb.codes, pb.theta, sigma=0.1,
                                   tol=tol, max_iter=50, block_size=int(pb.n * 0.05), seed=7,
                                   stabilized=True)
    assert 4 <= n_iter.value < 50
    assert n_iter.value == n_r


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("method", ["batched", "fast", "original"])
@pytest.mark.parametrize("flavor", ["harmony1", "harmony2"])
def test_correction(device, dtype, method, flavor):
    H = Hold(device)
    pb = Problem(n=2500, dtype=np.float64, B=4, K=10)
    Rr, Er, Or, _ = ho.init_clusters(pb.Z_norm, pb.Y0, sigma=0.1, Pr_b=pb.Pr_b, theta=pb.theta,
                                     cats=pb.codes, n_batches=pb.B, stabilized=True)
    X = pb.X
    dyn = flavor == "harmony2"
    lam_r = ho.compute_lambda_kb(Er, Or, pb.N_b, alpha=0.2, threshold=1e-5, ridge_lambda=1.0,
                                 dynamic_lambda=dyn)
    if method == "original":
        Zr = ho.correction_original(X, Rr, lam_r, pb.codes, pb.B)
    else:
        Zr, _ = ho.correction_batched(X, Rr, Or, lam_r, pb.codes, pb.B)
    tdt = _tdt(dtype)
    lib = _native.load()
    dm = _native.dims(pb.n, pb.d, pb.K, pb.B, int(pb.n * 0.05), np.dtype(dtype).itemsize)
    nb = lib.hz_workspace_bytes(ctypes.byref(dm))
    ws = torch.empty(nb, dtype=torch.uint8, device=device)
    Xt = _t(X.astype(dtype), device)
    Zh = torch.empty_like(Xt)
    Zn = torch.empty_like(Xt)
    lam = torch.empty((pb.B, pb.K), dtype=tdt, device=device)
    _native.check(lib.hz_correction(ctypes.byref(dm), _p(Xt), H(Rr.astype(dtype)),
                                    H(Or.astype(dtype)),
                                    H(Er.astype(dtype)),
                                    H(pb.N_b.astype(dtype)), H(pb.offsets),
                                    H(pb.cell_idx), _native.HZ_CORR[method], int(dyn),
                                    0.2, 1, 1e-5, 1.0, _p(lam), _p(Zh), _p(Zn), _p(ws), nb,
                                    _s(device)), "correction")
    tol = 1e-5 if dtype == np.float32 else 1e-10
    assert rel(Zh.cpu().numpy(), Zr) < tol
    assert rel(Zn.cpu().numpy(), ho.l2_normalize(Zr)) < tol
    if dtype == np.float64:
        np.testing.assert_array_equal(lam.cpu().numpy(), lam_r)


@pytest.mark.parametrize("flavor", ["harmony1", "harmony2"])
def test_harmonize_fp64_matches_oracle(device, flavor):
    pb = Problem(n=3000, d=20, B=3, K=15, dtype=np.float64)
    kw = dict(n_clusters=15, max_iter_harmony=3, max_iter_clustering=8, random_state=11,
              stabilized_penalty=flavor == "harmony2", dynamic_lambda=flavor == "harmony2")
    rec = {}
    Zg = harmonize(_t(pb.X, device), pb.obs, "batch", init_centroids=pb.Y0, record=rec, **kw)
    rec_r = {}
    Zr = ho.harmonize(pb.X, pb.codes, pb.B, Y0=pb.Y0, record=rec_r, **kw)
    assert rec["clustering_iters"] == rec_r["clustering_iters"]
    np.testing.assert_allclose(rec["objectives"], rec_r["objectives"], rtol=1e-9)
    assert rel(Zg.cpu().numpy(), Zr) < 1e-8


def test_harmonize_fp32_fixed_iters(device):
    """Plain FP32 path vs the oracle run in FP32 and FP64 at fixed iteration counts."""
    pb = Problem(n=4000, d=30, B=4, K=20, dtype=np.float32, n_types=8)
    kw = dict(n_clusters=20, max_iter_harmony=3, max_iter_clustering=5, random_state=5,
              tol_harmony=float("-inf"), tol_clustering=float("-inf"),
              stabilized_penalty=False, dynamic_lambda=False)
    Zg = harmonize(_t(pb.X, device), pb.obs, "batch", init_centroids=pb.Y0, **kw).cpu().numpy()
    Z64 = ho.harmonize(pb.X.astype(np.float64), pb.codes, pb.B, Y0=pb.Y0.astype(np.float64), **kw)
    Z32 = ho.harmonize(pb.X, pb.codes, pb.B, Y0=pb.Y0, **kw)
    e64 = rel(Zg, Z64)
    e32 = rel(Zg, Z32)
    assert e32 < 1e-4, (e32, e64)
    assert e64 < 1e-4, (e32, e64)


def test_harmony_integrate_dropin(device):
    X, obs = synthetic_pcs(3000, 20, 6, 3, seed=9)
    ad = SimpleAnnData(obs, {"X_pca": X})
    from rapids_singlecell_amd import pp

    pp.harmony_integrate(ad, "batch", dtype=np.float32, max_iter_harmony=2)
    out = ad.obsm["X_pca_harmony"]
    assert out.shape == X.shape and out.dtype == np.float32
    assert np.all(np.isfinite(out))


@pytest.mark.parametrize("flavor", ["harmony2", "harmony1"])
def test_harmony_integrate_dropin_matches_oracle(device, flavor):
    """The drop-in call with the reference's defaults (dtype float64, the
    reference's convergence rules) writes the oracle's Z_corr to
    obsm["X_pca_harmony"] when both start from the same centroids."""
    X, obs = synthetic_pcs(3000, 20, 6, 3, seed=9)
    ad = SimpleAnnData(obs, {"X_pca": X})
    from rapids_singlecell_amd import pp

    codes, cats, _ = ho.batch_codes(obs, "batch")
    X64 = X.astype(np.float64)
    Y0 = ho.l2_normalize(X64)[np.random.default_rng(4).choice(3000, 15, replace=False)]
    pp.harmony_integrate(ad, "batch", flavor=flavor, n_clusters=15, max_iter_clustering=30,
                         random_state=2, init_centroids=Y0)
    out = ad.obsm["X_pca_harmony"]
    assert out.dtype == np.float64
    h2 = flavor == "harmony2"
    Zr = ho.harmonize(X64, codes, len(cats), Y0=Y0, n_clusters=15, max_iter_clustering=30,
                      random_state=2, stabilized_penalty=h2, dynamic_lambda=h2)
    e = rel(out, Zr)
    print(f"\nharmony_integrate {flavor} fp64: rel err vs oracle = {e:.3e}", flush=True)
    assert e < 1e-8


def test_harmonize_tau_matches_oracle(device):
    """tau > 0 discounts theta per batch as theta * (1 - exp(-N_b/(K tau))**2)
    (_harmony/__init__.py:172-175), in the product path as in the oracle."""
    pb = Problem(n=3000, d=20, B=3, K=15, dtype=np.float64)
    # N_b / (K tau) ~ 0.07: the discount keeps about an eighth of theta; a
    # wider kernel (sigma 0.5) lets the penalty move R, so the discount shows
    kw = dict(n_clusters=15, max_iter_harmony=3, max_iter_clustering=8, random_state=11, tau=1000,
              sigma=0.5)
    rec, rec_r = {}, {}
    Zg = harmonize(_t(pb.X, device), pb.obs, "batch", init_centroids=pb.Y0, record=rec, **kw)
    Zr = ho.harmonize(pb.X, pb.codes, pb.B, Y0=pb.Y0, record=rec_r, **kw)
    assert rec["clustering_iters"] == rec_r["clustering_iters"]
    e = rel(Zg.cpu().numpy(), Zr)
    print(f"\nharmonize tau=1000 fp64: rel err vs oracle = {e:.3e}", flush=True)
    assert e < 1e-8
    # the discount changed the result (the test would pass vacuously otherwise)
    Zr0 = ho.harmonize(pb.X, pb.codes, pb.B, Y0=pb.Y0, **dict(kw, tau=0))
    assert rel(Zr0, Zr) > 1e-6


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_kmeans_matches_oracle(device, dtype):
    X, _ = synthetic_pcs(3000, 16, 5, 2, seed=21, center_scale=6.0)
    Z = ho.l2_normalize(X.astype(np.float64))
    from rapids_singlecell_amd.preprocessing._harmony import kmeans_centroids

    Yg = kmeans_centroids(_t(Z.astype(dtype), device), 8, 77).cpu().numpy()
    Yr = ho.kmeans_centroids(Z, 8, 77)
    tol = 1e-4 if dtype == np.float32 else 1e-8
    np.testing.assert_allclose(Yg, Yr, atol=tol)


@pytest.fixture
def tuning():
    """Restores the library's tuning knobs after a test changes them."""
    before = _native.get_tuning()
    yield _native.set_tuning
    _native.set_tuning(**before)


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("stabilized", [True, False])
def test_clustering_sweep_vs_per_block(device, tuning, precision, stabilized):
    """FP32 storage: the one-launch block sweep and the one-launch-per-block
    path run the same iteration; both agree with the FP64 oracle."""
    pb = Problem(n=12000, d=30, B=5, K=24, n_types=8, dtype=np.float32, dirichlet=True)
    lib = _native.load()
    dm = _native.dims(pb.n, pb.d, pb.K, pb.B, int(pb.n * 0.05), 4, precision)
    Z64 = pb.Z_norm.astype(np.float64)
    Rr, Er, Or, _ = ho.init_clusters(Z64, pb.Y0.astype(np.float64), sigma=0.1,
                                     Pr_b=pb.Pr_b.astype(np.float64),
                                     theta=pb.theta.astype(np.float64), cats=pb.codes,
                                     n_batches=pb.B, stabilized=stabilized)
    H = Hold(device)
    out = {}
    for no_sweep in (0, 1):
        tuning(no_sweep=no_sweep)
        # the layout (chunk size) follows the tuning: size the workspace under it
        nb = lib.hz_workspace_bytes(ctypes.byref(dm))
        assert nb > 0, lib.hz_last_error()
        ws = torch.empty(nb, dtype=torch.uint8, device=device)
        R = _t(Rr.astype(np.float32), device)
        E = _t(Er.astype(np.float32), device)
        O = _t(Or.astype(np.float32), device)
        last = torch.zeros(1, dtype=torch.float32, device=device)
        n_iter = ctypes.c_int(0)
        _native.check(lib.hz_clustering_loop(ctypes.byref(dm), H(pb.Z_norm), _p(R), _p(E), _p(O),
                                             H(pb.Pr_b.ravel()), H(pb.codes), H(pb.theta), 0.1,
                                             float("-inf"), 4, 99, int(stabilized), _p(last),
                                             ctypes.byref(n_iter), _p(ws), nb, _s(device)),
                      "loop")
        assert n_iter.value == 4
        out[no_sweep] = (R.cpu().numpy(), O.cpu().numpy(), E.cpu().numpy(), float(last.item()))
    last_r, _, _ = ho.clustering_loop(Z64, Rr, Er, Or, pb.Pr_b.astype(np.float64), pb.codes,
                                      pb.theta.astype(np.float64), sigma=0.1,
                                      tol=float("-inf"), max_iter=4,
                                      block_size=int(pb.n * 0.05), seed=99,
                                      stabilized=stabilized)
    for no_sweep in (0, 1):
        R, O, E, last = out[no_sweep]
        assert rel(R, Rr) < 1e-3, (no_sweep, rel(R, Rr))
        assert rel(O, Or) < 1e-4
        assert rel(E, Er) < 1e-4
        np.testing.assert_allclose(last, last_r, rtol=1e-4)
    assert rel(out[0][0], out[1][0]) < 1e-4


@pytest.mark.parametrize("dtype,precision", [(np.float32, "native"), (np.float32, "split_bf16"),
                                             (np.float64, "native")])
def test_block_update_launch_is_bitwise_the_same(device, tuning, dtype, precision):
    """One launch per block: the O/E/penalty update between blocks runs in the
    row launch's last workgroup or in a launch of its own spread over the
    clusters (the automatic choice above 512 levels x clusters, forced here).
    Every entry's arithmetic is the same, so the results are the same bits."""
    n, d, K, B = 20000, 20, 30, 6
    X, obs = synthetic_pcs(n, d, 6, B, seed=7)
    Xt = torch.as_tensor(X.astype(dtype), device=device)
    Zn = X / np.linalg.norm(X, axis=1, keepdims=True)
    Y0 = Zn[np.random.default_rng(7).choice(n, K, replace=False)].astype(dtype)
    out = {}
    for mode in (1, 2):
        tuning(no_sweep=1, block_update=mode)
        Z = harmonize(Xt, obs, "batch", n_clusters=K, init_centroids=Y0, random_state=0,
                      max_iter_harmony=2, max_iter_clustering=5, precision=precision)
        out[mode] = Z.cpu().numpy()
    assert np.isfinite(out[1]).all()
    np.testing.assert_array_equal(out[2], out[1])


@pytest.mark.parametrize("precision", ["native", "split_bf16"])
@pytest.mark.parametrize("K,d", [(50, 50), (75, 30), (64, 17), (33, 64), (130, 50)])
def test_harmonize_cluster_counts(device, precision, K, d):
    """Cluster counts that are not multiples of 4 (the block sweep stores R
    rows with 4-byte stores there, 16-byte stores otherwise), an odd PC count,
    and K > 128 (16 column tiles): agreement with the FP64 oracle at fixed
    iteration counts."""
    n, B = 12000, 5
    X, obs = synthetic_pcs(n, d, 12, B, seed=K + d)
    X = X.astype(np.float32)
    codes, cats, _ = ho.batch_codes(obs, "batch")
    Z64 = ho.l2_normalize(X.astype(np.float64))
    Y0 = Z64[np.random.default_rng(K).choice(n, K, replace=False)]
    kw = dict(n_clusters=K, max_iter_harmony=2, max_iter_clustering=3, random_state=3,
              tol_harmony=float("-inf"), tol_clustering=float("-inf"))
    Zg = harmonize(_t(X, device), obs, "batch", init_centroids=Y0.astype(np.float32),
                   precision=precision, **kw).cpu().numpy()
    Zr = ho.harmonize(X.astype(np.float64), codes, len(cats), Y0=Y0, **kw)
    assert np.all(np.isfinite(Zg))
    e = rel(Zg, Zr)
    print(f"\nK={K} d={d} {precision}: rel err vs FP64 oracle = {e:.3e}", flush=True)
    assert e < (1e-4 if precision == "native" else 1e-3), e


@pytest.mark.parametrize("precision", ["native", "split_bf16"])
@pytest.mark.parametrize("n,d,B", [(5395, 30, 7), (5836, 100, 25), (11227, 30, 28), (29263, 50, 32)])
def test_harmonize_shapes(device, precision, n, d, B):
    """Default-K jobs of the served-form distribution (LDS sizes around the
    64 KiB opt-in boundary, d=100, up to 32 batches) run end to end and agree
    with the FP64 oracle at fixed iteration counts."""
    X, obs = synthetic_pcs(n, d, 12, B, seed=n % 97)
    X = X.astype(np.float32)
    codes, cats, _ = ho.batch_codes(obs, "batch")
    K = max(int(min(100, n / 30)), 2)
    Z64 = ho.l2_normalize(X.astype(np.float64))
    Y0 = Z64[np.random.default_rng(5).choice(n, K, replace=False)]
    kw = dict(n_clusters=K, max_iter_harmony=2, max_iter_clustering=3, random_state=2,
              tol_harmony=float("-inf"), tol_clustering=float("-inf"))
    Zg = harmonize(_t(X, device), obs, "batch", init_centroids=Y0.astype(np.float32),
                   precision=precision, **kw).cpu().numpy()
    Zr = ho.harmonize(X.astype(np.float64), codes, len(cats), Y0=Y0, **kw)
    assert np.all(np.isfinite(Zg))
    e = rel(Zg, Zr)
    print(f"\nshape n={n} d={d} B={B} K={K} {precision}: rel err vs FP64 oracle = {e:.3e}",
          flush=True)
    assert e < (1e-4 if precision == "native" else 1e-3), e
