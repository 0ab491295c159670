"""GPU parity of the LDS segmented rank kernels (cs_rank, parameter_scan,
rolling_ic) through the C ABI, against the oracle. Ranks are compared
bitwise (NaN payload included); IC values exactly (the GPU and the oracle
both finalize exact integer sums)."""
import numpy as np
import pytest

import capi
from oracle import c_oracle, exact, np_arm

pytestmark = pytest.mark.gpu


def _panel(rng, T, N, nan_frac=0.02, decimals=2, inf=True):
    X = np.round(rng.standard_normal((T, N)), decimals).astype(np.float32)
    X[rng.random((T, N)) < nan_frac] = np.nan
    if inf and N > 4:
        X[0, 1] = np.inf
        X[min(1, T - 1), 2] = -np.inf
        X[0, 3] = -0.0
    return X


def _bits(a):
    return np.ascontiguousarray(a, dtype=np.float32).view(np.uint32)


SHAPES = [(1, 1), (1, 3), (3, 64), (5, 65), (7, 500), (4, 1000), (3, 5000), (2, 5003), (2, 8191)]


@pytest.mark.parametrize("T,N", SHAPES)
@pytest.mark.parametrize("desc", [False, True])
@pytest.mark.parametrize("masked", [False, True])
def test_cs_rank_bitwise(T, N, desc, masked):
    rng = np.random.default_rng(T * 1000 + N)
    X = _panel(rng, T, N)
    M = (rng.random((T, N)) < 0.8) if masked else None
    rc, got = capi.cs_rank(X, M, desc)
    assert rc == 0
    ref = c_oracle.rank_rows(X, M, desc)
    assert np.array_equal(_bits(got), _bits(ref))


def test_cs_rank_anchor_rows():
    # constant row, all-invalid row, dense ties, +-0 fold
    X = np.array([[2.0, 2.0, 2.0, 2.0],
                  [np.nan, np.inf, -np.inf, np.nan],
                  [1.0, -0.0, 0.0, 1.0],
                  [3.0, 3.0, 1.0, 3.0]], dtype=np.float32)
    for desc in (False, True):
        rc, got = capi.cs_rank(X, None, desc)
        assert rc == 0
        assert np.array_equal(_bits(got), _bits(np_arm.np_cs_rank(X, None, desc)))


@pytest.mark.parametrize("T,N", [(3, 3), (6, 777), (5, 5000), (2, 4096)])
def test_parameter_scan_four_groups(T, N):
    rng = np.random.default_rng(N)
    X = _panel(rng, T, N, decimals=1)
    M = rng.random((T, N)) < 0.7
    rc, outs, st, _, _ = capi.parameter_scan(X, M)
    assert rc == 0 and st == [0, 0, 0, 0]
    refs = np_arm.np_parameter_scan(X, M)
    for g in range(4):
        assert np.array_equal(_bits(outs[g]), _bits(refs[g])), g


def test_parameter_scan_subsets():
    rng = np.random.default_rng(7)
    X = _panel(rng, 4, 300, decimals=1)
    M = rng.random((4, 300)) < 0.6
    refs = np_arm.np_parameter_scan(X, M)
    for active in ([1, 0, 0, 0], [0, 0, 1, 0], [1, 0, 1, 0], [0, 1, 0, 1], [0, 0, 0, 1]):
        rc, outs, st, _, _ = capi.parameter_scan(X, M, active)
        assert rc == 0
        for g in range(4):
            if active[g]:
                assert st[g] == 0
                assert np.array_equal(_bits(outs[g]), _bits(refs[g])), (active, g)
            else:
                assert st[g] == -100


@pytest.mark.parametrize("T,N", [(1, 2), (5, 35), (20, 500), (8, 5000), (3, 3000)])
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_rolling_ic_exact(T, N, dtype):
    rng = np.random.default_rng(N + T)
    F = np.round(rng.standard_normal((T, N)), 2).astype(dtype)
    R = (0.3 * F + rng.standard_normal((T, N))).astype(dtype)
    F[rng.random((T, N)) < 0.02] = np.nan
    R[rng.random((T, N)) < 0.02] = np.inf
    fm = rng.random((T, N)) < 0.9
    rm = rng.random((T, N)) < 0.9
    mv = min(30, max(2, N // 2))
    rc, ic = capi.rolling_ic(F, R, fm, rm, mv)
    assert rc == 0
    ref = c_oracle.rolling_ic(F, R, fm, rm, mv)
    assert np.array_equal(np.isnan(ic), np.isnan(ref))
    assert np.array_equal(ic[~np.isnan(ic)], ref[~np.isnan(ref)])
    arm = np_arm.np_rolling_ic(F, R, fm, rm, mv)
    ok = ~np.isnan(arm)
    assert np.max(np.abs(ic[ok] - arm[ok]), initial=0.0) <= 1e-12


def test_rolling_ic_cpu_arm_config():
    """The CPU-arm workload (T=250, N=500, f32, min_valid=30): the device IC
    equals the C oracle bitwise and the numpy arm within 1e-12."""
    from bench import make_ic_cpu_inputs
    X, Y, M = make_ic_cpu_inputs()
    rc, ic = capi.rolling_ic(X, Y, M, None, 30)
    assert rc == 0
    assert np.array_equal(ic, c_oracle.rolling_ic(X, Y, M, None, 30))
    assert np.max(np.abs(ic - np_arm.np_rolling_ic(X, Y, M, None, 30))) <= 1e-12


def test_rolling_ic_ranks_and_degenerate():
    rng = np.random.default_rng(3)
    T, N = 6, 40
    F = np.round(rng.standard_normal((T, N)), 1)
    R = rng.standard_normal((T, N))
    F[1] = 0.5          # constant factor row -> NaN
    R[2] = 0.1          # constant return row -> NaN
    F[3, :36] = np.nan  # 4 valid < min_valid -> NaN
    rc, ic, rf, rr = capi.rolling_ic(F, R, None, None, 5, ranks=True)
    assert rc == 0
    assert np.isnan(ic[1]) and np.isnan(ic[2]) and np.isnan(ic[3])
    ref = c_oracle.rolling_ic(F, R, None, None, 5)
    assert np.array_equal(np.isnan(ic), np.isnan(ref))
    ok = ~np.isnan(ref)
    assert np.array_equal(ic[ok], ref[ok])
    valid = np.isfinite(F) & np.isfinite(R)
    for t in range(T):
        v = valid[t]
        er = np.zeros(N)
        er[v] = exact.stable_rank_row(F[t][v].astype(np.float32), np.ones(v.sum(), bool))
        # float64 ranks of float64 data: compare against argsort ranks
        o = np.argsort(F[t][v], kind="stable")
        er2 = np.zeros(v.sum())
        er2[o] = np.arange(1, v.sum() + 1)
        assert np.array_equal(rf[t][v], er2)
        assert np.all(rf[t][~v] == 0)
        o = np.argsort(R[t][v], kind="stable")
        er2[o] = np.arange(1, v.sum() + 1)
        assert np.array_equal(rr[t][v], er2)


def _extreme(rng, T, N, dtype):
    """Finite values spread over the whole exponent range (random bit
    patterns, both signs, subnormals included), with +-0, +-max, +-min
    normal, +-smallest subnormal, inf and NaN planted in every row."""
    if dtype == np.float32:
        bits = rng.integers(0, 2**32, (T, N), dtype=np.uint64).astype(np.uint32)
        X = bits.view(np.float32).copy()
    else:
        bits = rng.integers(0, 2**63, (T, N), dtype=np.uint64) * 2 + rng.integers(0, 2, (T, N), dtype=np.uint64)
        X = bits.view(np.float64).copy()
    X[~np.isfinite(X)] = 1.5  # keep the drawn part finite; specials below
    fi = np.finfo(dtype)
    specials = np.array([0.0, -0.0, fi.max, -fi.max, fi.tiny, -fi.tiny, fi.smallest_subnormal,
                         -fi.smallest_subnormal, np.inf, -np.inf, np.nan], dtype=dtype)
    for t in range(T):
        cols = rng.choice(N, size=len(specials), replace=False)
        X[t, cols] = specials
    return X


@pytest.mark.parametrize("N", [40, 700, 5000])
@pytest.mark.parametrize("desc", [False, True])
def test_cs_rank_whole_key_range(N, desc):
    rng = np.random.default_rng(N + desc)
    X = _extreme(rng, 4, N, np.float32)
    M = rng.random((4, N)) < 0.85
    for m in (None, M):
        rc, got = capi.cs_rank(X, m, desc)
        assert rc == 0
        assert np.array_equal(_bits(got), _bits(c_oracle.rank_rows(X, m, desc)))


def test_parameter_scan_whole_key_range():
    rng = np.random.default_rng(17)
    X = _extreme(rng, 3, 2100, np.float32)
    M = rng.random((3, 2100)) < 0.7
    rc, outs, st, _, _ = capi.parameter_scan(X, M)
    assert rc == 0 and st == [0, 0, 0, 0]
    refs = np_arm.np_parameter_scan(X, M)
    for g in range(4):
        assert np.array_equal(_bits(outs[g]), _bits(refs[g])), g


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_rolling_ic_whole_key_range(dtype):
    rng = np.random.default_rng(23)
    T, N = 6, 3000
    F = _extreme(rng, T, N, dtype)
    R = _extreme(rng, T, N, dtype)
    fm = rng.random((T, N)) < 0.9
    rm = rng.random((T, N)) < 0.9
    rc, ic = capi.rolling_ic(F, R, fm, rm, 30)
    assert rc == 0
    ref = c_oracle.rolling_ic(F, R, fm, rm, 30)
    assert np.array_equal(np.isnan(ic), np.isnan(ref))
    assert np.array_equal(ic[~np.isnan(ic)], ref[~np.isnan(ref)])

