"""GPU parity of rows beyond the register-resident rank kernels (more than
512*24 = 12288 columns, up to the 2^24 cap): cs_rank, parameter_scan and
rolling_ic go through a stable segmented sort of each row's (key, column)
pairs. Ranks are compared bitwise with the oracle and IC exactly, just
above the boundary and at the maximum row length; the device entry points
(strided planes, rank_out) and the allocator plan are checked too."""
import ctypes

import numpy as np
import pytest

import capi
from oracle import c_oracle, np_arm

pytestmark = pytest.mark.gpu

vp = ctypes.c_void_p


def _bits(a):
    return np.ascontiguousarray(a, dtype=np.float32).view(np.uint32)


def _row_panel(rng, T, N, dtype=np.float32):
    X = np.round(rng.standard_normal((T, N)), 2).astype(dtype)  # many ties
    X[rng.random((T, N)) < 0.02] = np.nan
    X[0, :4] = [np.inf, -np.inf, -0.0, 0.0]
    return X


def test_boundary_is_the_register_limit():
    assert capi.lib().fc_rank_max_n(4) == 12288


@pytest.mark.parametrize("T,N", [(3, 12288), (3, 12289), (2, 20000)])
@pytest.mark.parametrize("desc", [False, True])
@pytest.mark.parametrize("masked", [False, True])
def test_cs_rank_wide(T, N, desc, masked):
    rng = np.random.default_rng(N + 7 * desc + masked)
    X = _row_panel(rng, T, N)
    M = rng.random((T, N)) < 0.8 if masked else None
    rc, got = capi.cs_rank(X, M, desc)
    assert rc == 0
    assert np.array_equal(_bits(got), _bits(c_oracle.rank_rows(X, M, desc)))


def test_cs_rank_maximum_row():
    """One row of 2^24 columns, the reference's cap."""
    rng = np.random.default_rng(24)
    N = 1 << 24
    X = rng.standard_normal((1, N)).astype(np.float32)
    X[0, rng.integers(0, N, 1000)] = np.nan
    X[0, :1000] = np.round(X[0, :1000], 1)  # ties
    rc, got = capi.cs_rank(X, None, False)
    assert rc == 0
    assert np.array_equal(_bits(got), _bits(c_oracle.rank_rows(X, None, False)))
    assert capi.cs_rank(np.zeros((1, N + 1), np.float32))[0] == -2


def test_parameter_scan_wide():
    rng = np.random.default_rng(5)
    X = _row_panel(rng, 2, 13001)
    M = rng.random((2, 13001)) < 0.7
    rc, outs, st, _, _ = capi.parameter_scan(X, M)
    assert rc == 0 and st == [0, 0, 0, 0]
    refs = np_arm.np_parameter_scan(X, M)
    for g in range(4):
        assert np.array_equal(_bits(outs[g]), _bits(refs[g])), g


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_rolling_ic_wide(dtype):
    rng = np.random.default_rng(11)
    T, N = 4, 15000
    F = _row_panel(rng, T, N, dtype)
    R = (0.3 * F + rng.standard_normal((T, N))).astype(dtype)
    R[rng.random((T, N)) < 0.02] = np.inf
    F[1] = 0.5                 # constant factor row -> NaN
    fm = rng.random((T, N)) < 0.9
    rm = rng.random((T, N)) < 0.9
    rc, ic, rf, rr = capi.rolling_ic(F, R, fm, rm, 30, ranks=True)
    assert rc == 0
    ref = c_oracle.rolling_ic(F, R, fm, rm, 30)
    assert np.isnan(ic[1])
    assert np.array_equal(np.isnan(ic), np.isnan(ref))
    ok = ~np.isnan(ref)
    assert np.array_equal(ic[ok], ref[ok])
    valid = np.isfinite(F) & np.isfinite(R) & fm & rm
    for t in range(T):
        v = valid[t]
        for got, side in ((rf[t], F[t]), (rr[t], R[t])):
            o = np.argsort(side[v], kind="stable")
            want = np.zeros(v.sum())
            want[o] = np.arange(1, v.sum() + 1)
            assert np.array_equal(got[v], want)
            assert np.all(got[~v] == 0)


def test_rolling_ic_maximum_row():
    rng = np.random.default_rng(25)
    N = 1 << 24
    F = rng.standard_normal((1, N)).astype(np.float32)
    R = (0.1 * F + rng.standard_normal((1, N))).astype(np.float32)
    rc, ic = capi.rolling_ic(F, R, None, None, 30)
    assert rc == 0
    ref = c_oracle.rolling_ic(F, R, None, None, 30)
    assert np.array_equal(ic, ref)
    # an independent float64 Pearson of the ranks (the reference's CPU arm)
    assert abs(ic[0] - np_arm.np_rolling_ic(F, R, None, None, 30)[0]) <= 1e-12


def test_device_entry_points_wide():
    """fc_rank_scan_dev over strided (T, N, F) planes and fc_rank_ic_dev
    with rank_out, both beyond the register limit, against the host
    oracle per plane."""
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(9)
    T, N, F = 2, 12800, 3
    X = np.round(rng.standard_normal((T, N, F)), 2).astype(np.float32)
    M = rng.random((T, N)) < 0.85
    R = (0.2 * X[..., 0] + rng.standard_normal((T, N))).astype(np.float32)
    RM = rng.random((T, N)) < 0.95
    dev = "cuda:0"
    dX = torch.from_numpy(X).to(dev)
    dM = torch.from_numpy(M.view(np.uint8)).to(dev)
    dR = torch.from_numpy(R).to(dev)
    dRM = torch.from_numpy(RM.view(np.uint8)).to(dev)
    L = capi.lib()
    out = torch.full((F, T, N), -1.0, dtype=torch.float32, device=dev)
    rc = L.fc_rank_scan_dev(vp(dX.data_ptr()), F, T, N, 1, N * F, F, vp(dM.data_ptr()), 0,
                            vp(out.data_ptr()), None, None, None, None)
    assert rc == 0
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    for f in range(F):
        assert np.array_equal(_bits(got[f]), _bits(c_oracle.rank_rows(X[..., f], M))), f
    ic = torch.full((F, T), -1.0, dtype=torch.float64, device=dev)
    rk = torch.full((F, T, N), -1.0, dtype=torch.float32, device=dev)
    ws = capi.Workspace()
    rc = L.fc_rank_ic_dev(vp(dX.data_ptr()), 4, F, T, N, 1, N * F, F, vp(dM.data_ptr()), 0,
                          vp(dR.data_ptr()), vp(dRM.data_ptr()), 30, vp(ic.data_ptr()),
                          vp(rk.data_ptr()), ws.h, None)
    assert rc == 0
    torch.cuda.synchronize()
    icn, rkn = ic.cpu().numpy(), rk.cpu().numpy()
    for f in range(F):
        want = c_oracle.rolling_ic(np.ascontiguousarray(X[..., f]), R, M, RM, 30)
        assert np.array_equal(icn[f], want, equal_nan=True), f
        assert np.array_equal(_bits(rkn[f]), _bits(c_oracle.rank_rows(X[..., f], M))), f


def _hwm_of(call):
    L = capi.lib()
    s0 = capi.mem_stats()
    L.fc_mem_reset_peak()
    rc = call()
    s1 = capi.mem_stats()
    assert rc == 0, rc
    assert s1.live_bytes == s0.live_bytes
    return s1.peak_bytes - s0.live_bytes


def test_wide_rows_hwm_equals_plan():
    rng = np.random.default_rng(13)
    T, N = 2, 13000
    X = _row_panel(rng, T, N)
    M = rng.random((T, N)) < 0.9
    hwm = _hwm_of(lambda: capi.cs_rank(X, M)[0])
    assert hwm == capi.op_bytes(capi.OP_CS_RANK, T, N, flags=capi.PLAN_MASK)
    R = rng.standard_normal((T, N)).astype(np.float32)
    hwm = _hwm_of(lambda: capi.rolling_ic(X, R, M, None, 30)[0])
    assert hwm == capi.op_bytes(capi.OP_ROLLING_IC, T, N, flags=capi.PLAN_MASK)
