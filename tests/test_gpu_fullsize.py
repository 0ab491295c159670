"""Parity at the benchmark shapes, through the C ABI, against the oracle.

The reference gated its correlations on every pair of the full-size panels
(benchmarks/results/corpus_parity_v1.md:17-19: all 12,502,500 stock pairs at
N=5000, all 78 factor pairs at F=12) and its ranks bitwise on the full
corpus (benchmarks/results/phase4_bench_v1.md:20-21). Here:

- stock_corr, masked one-factor market model, T=1218, N in {2000, 5000}
  (bench.make_stock_inputs): every cell against the restated numpy arm on
  the float64 upcast of the float32 panel; contract tier within 1e-12, fp32
  and bf16x3 within 1e-5; NaN pattern equal on every cell.
- factor_corr, LKJ(2)-correlated float32 panel with a Bernoulli(0.9) mask,
  T=1218, N=5000, F=12 in all tiers; F=128 at T=256 (the float64 oracle at
  T=1218 needs ~30 GB of host memory and ~100 s).
- The batched rank scan (fc_rank_scan_dev) of T=1218, N=5000, F=12 planes
  (bench.make_rank_inputs: uneven per-date valid fractions, ~2 % ties), all
  four groups, in the plane-major (F,T,N) and factor-fastest (T,N,F)
  layouts: every plane of every group bitwise against the C oracle; and the
  same at N in {300, 12288} (one and the largest register-resident row).
"""
import ctypes

import numpy as np
import pytest

import bench
import capi
from oracle import c_oracle, np_arm

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

TOL_CONTRACT = 1e-12
TOL_FAST = 1e-5


def corr_match(got, exp, tol):
    assert got.shape == exp.shape
    nan_g, nan_e = np.isnan(got), np.isnan(exp)
    assert np.array_equal(nan_g, nan_e), (
        f"NaN pattern differs in {int((nan_g != nan_e).sum())} cells")
    ok = ~nan_e
    d = float(np.abs(got[ok] - exp[ok]).max()) if ok.any() else 0.0
    assert d <= tol, f"max |dr| = {d:.3e} > {tol}"
    return d


# ---- stock_corr (T=1218, N=2000 / 5000) ------------------------------------
@pytest.fixture(scope="module", params=[2000, 5000], ids=["N2000", "N5000"])
def stock_case(request):
    N = request.param
    X, M = bench.make_stock_inputs(N)
    ref = np_arm.np_stock_corr(X.astype(np.float64), M)
    return N, X, M, ref


@pytest.mark.parametrize("tier", [0, 1, 2], ids=["contract", "fp32", "bf16x3"])
def test_stock_corr_full_size_vs_oracle(stock_case, tier):
    N, X, M, ref = stock_case
    print(f"stock_corr T=1218 N={N} tier={tier}", flush=True)
    rc, out, (path, nfb) = capi.stock_corr(X, M, tier)
    assert rc == 0
    assert path == 0  # listing gaps and holes: the general (masked) path
    assert np.array_equal(out.view(np.uint64), out.T.view(np.uint64)), "mirror must be bitwise"
    err = corr_match(out, ref, TOL_CONTRACT if tier == 0 else TOL_FAST)
    print(f"  max|dr| vs oracle {err:.3e}, exact-pass cells {nfb}", flush=True)


def test_stock_corr_full_size_fully_valid_path():
    """The same market-model values with every cell valid: the fully valid
    path (selected_path 1) against the oracle at N=5000."""
    X, _ = bench.make_stock_inputs(5000)
    ref = np_arm.np_stock_corr(X.astype(np.float64), None)
    for tier in (0, 1, 2):
        rc, out, (path, nfb) = capi.stock_corr(X, None, tier)
        assert rc == 0 and path == 1
        err = corr_match(out, ref, TOL_CONTRACT if tier == 0 else TOL_FAST)
        print(f"stock_corr fully valid N=5000 tier={tier}: max|dr| {err:.3e}, exact-pass {nfb}",
              flush=True)


# ---- factor_corr (T=1218, N=5000, F=12; F=128 at T=256) --------------------
def factor_panel(T, N, F, seed=2):
    """x = L z, L = chol(C), C ~ LKJ(eta=2) by the onion method
    (bench.lkj_onion), z ~ N(0, I_F) float32; mask Bernoulli(0.9) on (T,N)."""
    rng = np.random.default_rng(seed)
    C = bench.lkj_onion(F, 2.0, rng)
    L = np.linalg.cholesky(C).astype(np.float32)
    Z = rng.standard_normal((T * N, F), dtype=np.float32)
    X = np.ascontiguousarray((Z @ L.T).reshape(T, N, F))
    M = rng.random((T, N), dtype=np.float32) < 0.9
    return X, M


@pytest.fixture(scope="module", params=[(1218, 12), (256, 128)], ids=["T1218F12", "T256F128"])
def factor_case(request):
    T, F = request.param
    X, M = factor_panel(T, 5000, F)
    ref = np_arm.np_factor_corr(X, M)
    return T, F, X, M, ref


@pytest.mark.parametrize("tier", [0, 1, 2], ids=["contract", "fp32", "bf16x3"])
def test_factor_corr_full_size_vs_oracle(factor_case, tier):
    T, F, X, M, ref = factor_case
    print(f"factor_corr T={T} N=5000 F={F} tier={tier}", flush=True)
    rc, out = capi.factor_corr(X, M, tier)
    assert rc == 0
    assert np.array_equal(out.view(np.uint64), out.T.view(np.uint64))
    err = corr_match(out, ref, TOL_CONTRACT if tier == 0 else TOL_FAST)
    print(f"  max|dr| vs oracle {err:.3e}", flush=True)


# ---- batched rank scan: F planes x 4 groups in one launch ------------------
def _scan_dev(dX, F, T, N, strides, dM):
    dev = dX.device
    outs = [torch.empty((F, T, N), dtype=torch.float32, device=dev) for _ in range(4)]
    vp = ctypes.c_void_p
    ps, rs, es = strides
    rc = capi.lib().fc_rank_scan_dev(vp(dX.data_ptr()), F, T, N, ps, rs, es, vp(dM.data_ptr()), 0,
                                     *[vp(o.data_ptr()) for o in outs], None)
    assert rc == 0, rc
    torch.cuda.synchronize()
    return [o.cpu().numpy() for o in outs]


def _scan_oracle(X, M):
    """[group][plane] ranks: asc-masked, asc-unmasked, desc-masked, desc-unmasked."""
    return [[c_oracle.rank_rows(X[f], m, d) for f in range(X.shape[0])]
            for (m, d) in ((M, False), (None, False), (M, True), (None, True))]


def _check_scan(got, ref, label):
    for g in range(4):
        for f in range(len(ref[g])):
            assert np.array_equal(got[g][f].view(np.uint32), ref[g][f].view(np.uint32)), (
                f"{label}: group {g} plane {f} differs")


def _scan_both_layouts(X, M, label):
    """X (F,T,N) float32, M (T,N) bool."""
    F, T, N = X.shape
    ref = _scan_oracle(X, M)
    dev = torch.device("cuda:0")
    dM = torch.from_numpy(np.ascontiguousarray(M).view(np.uint8)).to(dev)
    print(f"{label}: (F,T,N) layout", flush=True)
    dX = torch.from_numpy(X).to(dev)
    _check_scan(_scan_dev(dX, F, T, N, (T * N, N, 1), dM), ref, label + " (F,T,N)")
    del dX
    print(f"{label}: (T,N,F) layout", flush=True)
    dXt = torch.from_numpy(np.ascontiguousarray(X.transpose(1, 2, 0))).to(dev)
    _check_scan(_scan_dev(dXt, F, T, N, (1, N * F, F), dM), ref, label + " (T,N,F)")


def test_rank_scan_batched_full_size():
    """T=1218, N=5000, F=12, per-date valid fraction U(0.6,1), values
    rounded to 3 decimals: the benchmark's batched parameter_scan."""
    X, M, _, _ = bench.make_rank_inputs()
    _scan_both_layouts(X, M, "scan T=1218 N=5000 F=12")


@pytest.mark.parametrize("N", [300, 12288])
def test_rank_scan_batched_row_lengths(N):
    rng = np.random.default_rng(N)
    T, F = 97, 12
    p = rng.uniform(0.6, 1.0, size=T)
    M = rng.random((T, N)) < p[:, None]
    X = np.round(rng.standard_normal((F, T, N)), 3).astype(np.float32)
    X[1, 3, : N // 3] = np.nan     # invalid cells in unmasked groups too
    X[2, 5, :] = 0.5               # a constant row
    X[3, 7, ::2] = -0.0            # +-0 ties
    X[3, 7, 1::2] = 0.0
    _scan_both_layouts(X, M, f"scan T={T} N={N} F={F}")
