"""GPU parity of stock_corr / factor_corr through the C ABI.

Contract tier: |dr| <= 1e-12 against the per-pair np.corrcoef oracle
(tests/fixtures/corr_oracle_v1.py semantics) and against the restated numpy
arm, NaN pattern equal; on the HG-2 bias inputs against serial Kahan.
fp32 / bf16x3 tiers: max |dr| <= 1e-5 against the FP64 oracle, NaN pattern
equal, counts exact (a count error moves r far beyond 1e-5)."""
import json
import pathlib

import numpy as np
import pytest

import capi
from oracle import exact, np_arm

pytestmark = pytest.mark.gpu

GOLD = pathlib.Path(__file__).resolve().parent / "golden"
TOL_CONTRACT = 1e-12
TOL_FAST = 1e-5


def corr_match(got, exp, tol):
    got = np.asarray(got, dtype=np.float64)
    exp = np.asarray(exp, dtype=np.float64)
    assert np.array_equal(np.isnan(got), np.isnan(exp)), (
        f"NaN pattern differs at {np.argwhere(np.isnan(got) != np.isnan(exp))[:5]}")
    ok = ~np.isnan(exp)
    d = np.abs(got[ok] - exp[ok]) if ok.any() else np.zeros(1)
    assert d.max() <= tol, f"max |dr| = {d.max():.3e} > {tol}"
    return float(d.max())


def _corpus():
    z = np.load(GOLD / "corr_corpus_v1.npz", allow_pickle=False)
    m = json.loads((GOLD / "corr_corpus_v1.manifest.json").read_text())
    for c in m["cases"]:
        nm = c["id"]
        a, b = z[nm + "_a"], z[nm + "_b"]
        ma = z[nm + "_mask_a"] if nm + "_mask_a" in z.files else None
        mb = z[nm + "_mask_b"] if nm + "_mask_b" in z.files else None
        yield c, a, b, ma, mb


@pytest.mark.parametrize("op", ["stock", "factor"])
def test_corr_corpus_contract(op):
    for c, a, b, ma, mb in _corpus():
        v = np.isfinite(a) & np.isfinite(b)
        if ma is not None:
            v &= ma
        if mb is not None:
            v &= mb
        if c["branch"] == "kahan":
            exp = exact.serial_kahan_corr(a[v], b[v])   # HG-2 clause
        else:
            exp = float("nan") if c["expected"] is None else float(c["expected"])
        if op == "stock":
            X = np.ascontiguousarray(np.column_stack([a, b]))
            M = None
            if ma is not None or mb is not None:
                ones = np.ones(a.shape, bool)
                M = np.column_stack([ma if ma is not None else ones, mb if mb is not None else ones])
            if M is not None and (ma is not None) != (mb is not None):
                pass
            rc, out, _ = capi.stock_corr(X, M, 0)
        else:
            if ma is not None and mb is not None and not np.array_equal(ma, mb):
                continue  # factor_corr has one (T,N) mask for all factors
            F3 = np.ascontiguousarray(np.stack([a, b], axis=1).reshape(1, -1, 2))
            m = None if ma is None else ma.reshape(1, -1)
            rc, out = capi.factor_corr(F3, m, 0)
        assert rc == 0, (c["id"], rc)
        got = out[0, 1]
        assert out[0, 1] == out[1, 0] or (np.isnan(out[0, 1]) and np.isnan(out[1, 0]))
        if np.isnan(exp):
            assert np.isnan(got), (c["id"], got)
        else:
            assert abs(got - exp) <= TOL_CONTRACT, (c["id"], got, exp)


def _panel(rng, T, N, nan_frac=0.01, mask_p=0.9):
    X = rng.standard_normal((T, N)) * rng.uniform(0.5, 3.0, N) + rng.uniform(-2, 2, N)
    X[rng.random((T, N)) < nan_frac] = np.nan
    M = rng.random((T, N)) < mask_p
    if N > 6:
        X[:, 3] = 0.25          # constant column
        M[:, 5] = False         # all-invalid column
        X[: T // 2, 6] = np.inf  # half of a column non-finite
    return X, M


@pytest.mark.parametrize("T,N", [(5, 3), (40, 17), (300, 150), (130, 70)])
def test_stock_corr_contract_vs_oracles(T, N):
    rng = np.random.default_rng(T * 7 + N)
    X, M = _panel(rng, T, N)
    rc, out, st = capi.stock_corr(X, M, 0)
    assert rc == 0
    assert np.array_equal(out.view(np.uint64), out.T.view(np.uint64)), "mirror must be bitwise"
    ref = exact.corr_matrix_pairwise(X, M)
    corr_match(out, ref, TOL_CONTRACT)
    corr_match(out, np_arm.np_stock_corr(X, M), TOL_CONTRACT)


@pytest.mark.parametrize("tier", [1, 2])
@pytest.mark.parametrize("T,N", [(40, 17), (300, 150), (700, 260),
                                 # two T stages; N a multiple of the 128-row tile; 5 row blocks
                                 (3, 300), (129, 256), (65, 520)])
def test_stock_corr_fast_tiers(tier, T, N):
    rng = np.random.default_rng(N)
    X, M = _panel(rng, T, N)
    X32 = X.astype(np.float32)
    rc, out, st = capi.stock_corr(X32, M, tier)
    assert rc == 0
    ref = exact.corr_matrix_pairwise(X32.astype(np.float64), M)
    corr_match(out, ref, TOL_FAST)


def _market(rng, T, N):
    beta = rng.uniform(0.5, 1.5, N)
    mkt = rng.normal(0.0, 0.01, T)
    X = (mkt[:, None] * beta + 0.02 * rng.standard_t(4, (T, N))).astype(np.float32)
    start = rng.integers(0, min(600, T // 2) + 1, N)
    M = (np.arange(T)[:, None] >= start[None, :]) & ~(rng.random((T, N)) < 0.01)
    return X, M


@pytest.mark.parametrize("tier", [0, 1, 2])
def test_stock_corr_market_model(tier):
    rng = np.random.default_rng(3)
    X, M = _market(rng, 1218, 333)
    rc, out, st = capi.stock_corr(X, M, tier)
    assert rc == 0
    ref = np_arm.np_stock_corr(X.astype(np.float64), M)
    corr_match(out, ref, TOL_CONTRACT if tier == 0 else TOL_FAST)


def test_stock_corr_f32_equals_f64_upcast():
    rng = np.random.default_rng(11)
    X, M = _market(rng, 200, 90)
    rc1, a, _ = capi.stock_corr(X, M, 0)
    rc2, b, _ = capi.stock_corr(X.astype(np.float64), M, 0)
    assert rc1 == 0 and rc2 == 0
    assert np.array_equal(a, b, equal_nan=True)


def test_stock_corr_fully_valid_path():
    rng = np.random.default_rng(5)
    X = rng.standard_normal((257, 130)).astype(np.float32)
    for tier in (0, 1, 2):
        rc, out, st = capi.stock_corr(X, None, tier)
        assert rc == 0 and st[0] == 1
        ref = np.corrcoef(X.astype(np.float64).T)
        corr_match(out, ref, TOL_CONTRACT if tier == 0 else TOL_FAST)


def _adversarial(kind, rng, T, N):
    z = rng.standard_normal((T, N))
    if kind == "rounded":      # few distinct values: bf16 split errors repeat
        X = np.round(z, 2)
    elif kind == "digits":     # integers 0..9
        X = rng.integers(0, 10, (T, N)).astype(np.float64)
    elif kind == "lognormal":  # heavy right tail
        X = np.exp(2.0 * z)
    elif kind == "offset":     # large common offset, unit spread
        X = 1e4 + z
    else:                      # columns on scales 1e-3 .. 1e3
        X = z * 10.0 ** rng.uniform(-3, 3, N)[None, :]
    M = rng.random((T, N)) < 0.9
    return X.astype(np.float32), M


@pytest.mark.parametrize("kind", ["rounded", "digits", "lognormal", "offset", "scales"])
@pytest.mark.parametrize("tier", [1, 2])
def test_stock_corr_fast_tiers_adversarial_values(kind, tier):
    rng = np.random.default_rng(len(kind))
    X, M = _adversarial(kind, rng, 400, 120)
    rc, out, st = capi.stock_corr(X, M, tier)
    assert rc == 0
    ref = exact.corr_matrix_pairwise(X.astype(np.float64), M)
    corr_match(out, ref, TOL_FAST)


@pytest.mark.parametrize("T", [3, 9])
def test_stock_corr_fast_tiers_few_rows_fully_valid(T):
    """Fully valid panels of a few rows: every pair has n = T. bf16x3 sends
    such cells (n < 16) to the exact pass, so both fast tiers stay within
    1e-5 of the FP64 Pearson."""
    rng = np.random.default_rng(T)
    X = rng.standard_normal((T, 200)).astype(np.float32)
    for tier in (1, 2):
        rc, out, st = capi.stock_corr(X, None, tier)
        assert rc == 0 and st[0] == 1
        ref = np.corrcoef(X.astype(np.float64).T)
        corr_match(out, ref, TOL_FAST)


def test_stock_corr_domain_error():
    X = np.array([[1e151, 1.0], [2.0, 3.0], [4.0, 5.0]])
    rc, _, _ = capi.stock_corr(X, None, 0)
    assert rc == -4


def _fpanel(rng, T, N, F, nan=True):
    L = np.linalg.cholesky(exact_lkj(rng, F))
    Z = rng.standard_normal((T * N, F)) @ L.T
    F3 = Z.reshape(T, N, F)
    M = rng.random((T, N)) < 0.9
    if nan:
        # a masked non-finite cell: the pair-valid sets differ per pair and
        # the general six-sum path runs
        M[0, 0] = True
        F3[0, 0, 0] = np.nan
    return F3, M


def exact_lkj(rng, F):
    A = rng.standard_normal((F, 2 * F))
    C = A @ A.T
    d = np.sqrt(np.diag(C))
    return C / d[:, None] / d[None, :]


def factor_oracle(F3, M):
    T, N, F = F3.shape
    X = F3.reshape(T * N, F)
    m = None if M is None else np.repeat(M.reshape(-1, 1), F, axis=1)
    return exact.corr_matrix_pairwise(X, m)


@pytest.mark.parametrize("T,N,F", [(60, 15, 4), (50, 40, 12), (20, 30, 40), (10, 25, 128)])
@pytest.mark.parametrize("tier", [0, 1, 2])
@pytest.mark.parametrize("nan", [True, False])
def test_factor_corr_vs_oracle(T, N, F, tier, nan):
    rng = np.random.default_rng(T + N + F)
    F3, M = _fpanel(rng, T, N, F, nan)
    src = F3 if tier == 0 else F3.astype(np.float32)
    rc, out = capi.factor_corr(np.ascontiguousarray(src), M, tier)
    assert rc == 0
    assert np.array_equal(out.view(np.uint64), out.T.view(np.uint64))
    ref = factor_oracle(src.astype(np.float64), M)
    corr_match(out, ref, TOL_CONTRACT if tier == 0 else TOL_FAST)


@pytest.mark.parametrize("kind", ["rounded", "digits", "lognormal", "offset", "scales"])
@pytest.mark.parametrize("tier", [0, 1, 2])
def test_factor_corr_adversarial_values(kind, tier):
    """The adversarial value sets of the stock_corr test, pooled over
    (t, i) rows of a (T, N, F) panel (factor f = column f of a (T*N, F)
    draw)."""
    rng = np.random.default_rng(7 + len(kind))
    T, N, F = 60, 50, 10
    X, _ = _adversarial(kind, rng, T * N, F)
    F3 = np.ascontiguousarray(X.reshape(T, N, F))
    M = rng.random((T, N)) < 0.9
    rc, out = capi.factor_corr(F3, M, tier)
    assert rc == 0
    ref = factor_oracle(F3.astype(np.float64), M)
    corr_match(out, ref, TOL_CONTRACT if tier == 0 else TOL_FAST)


def test_factor_corr_f32_bitwise_equals_f64():
    rng = np.random.default_rng(9)
    F3, M = _fpanel(rng, 40, 12, 3)
    F32 = F3.astype(np.float32)
    rc1, a = capi.factor_corr(F32, M, 0)
    rc2, b = capi.factor_corr(F32.astype(np.float64), M, 0)
    assert rc1 == 0 and rc2 == 0
    assert np.array_equal(a, b, equal_nan=True)
    allfalse = np.zeros((40, 12), bool)
    rc, c = capi.factor_corr(F32, allfalse, 0)
    assert rc == 0 and np.isnan(c).all()


# every factor-block count of the uniform (all-finite) path: F from 1 to 128
# in 16-wide (contract, and F <= 16 in the fast tiers) and 32-wide blocks
@pytest.mark.parametrize("F", [1, 3, 16, 17, 31, 33, 48, 64, 65, 80, 96, 97, 112, 127, 128])
@pytest.mark.parametrize("tier", [0, 1, 2])
def test_factor_corr_uniform_path_block_counts(F, tier):
    rng = np.random.default_rng(1000 + F)
    T, N = 37, 41   # R = 1517 rows: a ragged last iteration in every plan
    F3, M = _fpanel(rng, T, N, F, nan=False)
    src = F3 if tier == 0 else F3.astype(np.float32)
    rc, out = capi.factor_corr(np.ascontiguousarray(src), M, tier)
    assert rc == 0
    assert np.array_equal(out.view(np.uint64), out.T.view(np.uint64))
    ref = factor_oracle(src.astype(np.float64), M)
    corr_match(out, ref, TOL_CONTRACT if tier == 0 else TOL_FAST)


@pytest.mark.parametrize("tier", [0, 1, 2])
def test_factor_corr_degenerate_columns(tier):
    """Constant, near-duplicate and sign-flipped factors: the diagonal of a
    constant factor is NaN, |r| = 1 pairs come out exact, a factor with a
    masked NaN switches to the general path."""
    rng = np.random.default_rng(77)
    T, N, F = 30, 50, 6
    X = rng.standard_normal((T, N, F))
    X[..., 1] = 2.5                      # constant column
    X[..., 2] = 3.0 * X[..., 0] + 1.0    # r = +1 with factor 0
    X[..., 3] = -X[..., 0]               # r = -1 with factor 0
    M = rng.random((T, N)) < 0.8
    for nan in (False, True):
        Y = X.copy()
        if nan:
            M2 = M.copy()
            M2[3, 4] = True
            Y[3, 4, 5] = np.inf
        else:
            M2 = M
        src = Y if tier == 0 else Y.astype(np.float32)
        rc, out = capi.factor_corr(np.ascontiguousarray(src), M2, tier)
        assert rc == 0
        ref = factor_oracle(src.astype(np.float64), M2)
        corr_match(out, ref, TOL_CONTRACT if tier == 0 else TOL_FAST)
        assert np.isnan(out[1, 1]) and np.isnan(out[0, 1])
        assert abs(out[0, 2] - 1.0) <= (TOL_CONTRACT if tier == 0 else TOL_FAST)
        assert abs(out[0, 3] + 1.0) <= (TOL_CONTRACT if tier == 0 else TOL_FAST)


@pytest.mark.parametrize("F", [12, 40])
@pytest.mark.parametrize("tier", [0, 1, 2])
def test_factor_corr_nonfinite_and_extreme_cells(F, tier):
    """How the Gram pass meets non-finite and extreme cells: a masked +inf
    and -inf in different factors must send the call to the general path
    (the pass sees them through its column sums); non-finite cells outside
    the mask must change nothing; a masked finite outlier near the float
    maximum must still give the oracle's matrix. F=12 is one 16-wide block,
    F=40 several blocks with idle lanes in the last."""
    rng = np.random.default_rng(500 + F)
    T, N = 30, 50
    F3, M = _fpanel(rng, T, N, F, nan=False)
    F3 = F3.astype(np.float32)
    cases = []
    Y = F3.copy()
    M1 = M.copy()
    M1[2, 3] = M1[5, 7] = True
    Y[2, 3, 1] = np.inf
    Y[5, 7, F - 1] = -np.inf
    cases.append((Y, M1))
    Y = F3.copy()
    M2 = M.copy()
    M2[4, 9] = M2[6, 1] = False
    Y[4, 9, 0] = np.nan
    Y[6, 1, F - 2] = np.inf
    cases.append((Y, M2))
    Y = F3.copy()
    M3 = M.copy()
    M3[7, 11] = True
    Y[7, 11, 3] = np.float32(3.0e38)
    cases.append((Y, M3))
    for Y, Mk in cases:
        src = np.ascontiguousarray(Y.astype(np.float64) if tier == 0 else Y)
        rc, out = capi.factor_corr(src, Mk, tier)
        assert rc == 0
        assert np.array_equal(out.view(np.uint64), out.T.view(np.uint64))
        ref = factor_oracle(src.astype(np.float64), Mk)
        corr_match(out, ref, TOL_CONTRACT if tier == 0 else TOL_FAST)
    # float input in the contract tier takes the late-scaled Gram pass
    for Y, Mk in cases:
        rc, out = capi.factor_corr(np.ascontiguousarray(Y), Mk, 0)
        assert rc == 0
        corr_match(out, factor_oracle(Y.astype(np.float64), Mk), TOL_CONTRACT)


def test_factor_corr_large_uniform_bitwise_repeatable():
    """Determinism: two calls on the same 1.2M-row panel agree bitwise in every
    tier (fixed row partition and fixed reduction order, no float atomics)."""
    rng = np.random.default_rng(5)
    T, N, F = 300, 4000, 12
    F3 = rng.standard_normal((T, N, F)).astype(np.float32)
    M = rng.random((T, N)) < 0.9
    for tier in (0, 1, 2):
        rc1, a = capi.factor_corr(F3, M, tier)
        rc2, b = capi.factor_corr(F3, M, tier)
        assert rc1 == 0 and rc2 == 0
        assert np.array_equal(a.view(np.uint64), b.view(np.uint64))
