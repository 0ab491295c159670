"""Contract-tier hardening and binding diagnostics on the GPU.

- HG-2 var-underflow class (tests/test_adapter_v1.py:749-783 of the
  reference): five in-domain values at the 1e-140 scale, stock_corr and
  factor_corr on backend='cuda' (contract tier) within 1e-12 of the serial
  Kahan correlation; the same through the device-resident route.
- stock_corr_f64(return_stats=True) through the pybind binding
  (src/factor_cuda_pybind.cpp:240-278; StockCorrRunStats,
  src/stock_corr.cuh:47-50): selected_path 0 on a masked panel and 1 on a
  fully valid one; fallback_count > 0 on the HG-2 bias panels, whose cells
  must take the exact pass, and equal to the C ABI's count.
- rolling_ic_f64(return_ranks=True) through the binding: the ordinal ranks
  of both sides on the pairwise-valid set bitwise against the C oracle,
  0 outside that set, and the IC equal to the C oracle's integer IC.
- Two threads on two torch streams sharing the adapter's device workspace:
  each result equals the C oracle (the cache lock is held until the
  call's kernels have drained).
"""
import json
import pathlib
import threading

import numpy as np
import pytest

import capi
import fc
from fc import _util as u
from oracle import c_oracle, exact, np_arm

pytestmark = pytest.mark.gpu

GOLD = pathlib.Path(__file__).resolve().parent / "golden"

XV = np.array([1e-140, 2e-140, 3e-140, 4e-140, 5e-140], dtype=np.float64)
YV = np.array([5e-140, 4e-140, 3e-140, 2e-140, 1e-140], dtype=np.float64)


def test_stock_corr_var_underflow_hg2():
    X = np.column_stack([XV, YV])
    ref = exact.serial_kahan_corr(XV, YV)
    got = float(np.asarray(fc.stock_corr(X, None, backend="cuda"))[0, 1])
    assert abs(got - ref) <= 1e-12, (got, ref)
    t = u._torch
    got_d = fc.stock_corr(t.from_numpy(X).to("cuda:0"), None, backend="cuda")
    assert abs(float(got_d[0, 1].item()) - ref) <= 1e-12


def test_factor_corr_var_underflow_hg2():
    F3 = np.stack([XV, YV], axis=1).reshape(1, -1, 2)
    ref = exact.serial_kahan_corr(XV, YV)
    got = float(np.asarray(fc.factor_corr(F3, None, backend="cuda"))[0, 1])
    assert abs(got - ref) <= 1e-12, (got, ref)
    t = u._torch
    got_d = fc.factor_corr(t.from_numpy(np.ascontiguousarray(F3)).to("cuda:0"), None,
                           backend="cuda")
    assert abs(float(got_d[0, 1].item()) - ref) <= 1e-12


def _bias_cases():
    z = np.load(GOLD / "corr_corpus_v1.npz", allow_pickle=False)
    m = json.loads((GOLD / "corr_corpus_v1.manifest.json").read_text())
    names = [c["id"] for c in m["cases"] if c["branch"] == "kahan"]
    assert {"bias_1e12", "bias_1e15"} <= set(names)
    return [(nm, z[nm + "_a"], z[nm + "_b"]) for nm in names]


def test_stock_corr_return_stats_paths():
    fcb = u.fcb()
    rng = np.random.default_rng(21)
    X = rng.standard_normal((300, 40))
    M = rng.random((300, 40)) < 0.9
    out, st = fcb.stock_corr_f64(X, M, return_stats=True)
    assert set(st) == {"selected_path", "fallback_count"}
    assert st["selected_path"] == 0
    out_v, st_v = fcb.stock_corr_f64(X, None, return_stats=True)
    assert st_v["selected_path"] == 1
    # return_stats does not change the matrix
    assert np.array_equal(out, fcb.stock_corr_f64(X, M), equal_nan=True)
    assert np.array_equal(out_v, fcb.stock_corr_f64(X, None), equal_nan=True)
    # an all-true mask is the fully valid path too (no invalid cell)
    _, st_t = fcb.stock_corr_f64(X, np.ones_like(M), return_stats=True)
    assert st_t["selected_path"] == 1


def test_stock_corr_return_stats_fallback_on_bias_panels():
    fcb = u.fcb()
    for name, a, b in _bias_cases():
        v = np.isfinite(a) & np.isfinite(b)
        X = np.column_stack([a, b])
        M = np.column_stack([v, v])
        out, st = fcb.stock_corr_f64(X, M, return_stats=True)
        ref = exact.serial_kahan_corr(a[v], b[v])
        assert abs(out[0, 1] - ref) <= 1e-12, (name, out[0, 1], ref)
        assert st["fallback_count"] > 0, (name, st)
        rc, out_c, (path, nfb) = capi.stock_corr(X, M, 0)
        assert rc == 0 and nfb == st["fallback_count"] and path == st["selected_path"]
        assert np.array_equal(out_c, out, equal_nan=True)
    # a well-conditioned panel: only the cells the cancellation gates flag
    # take the exact pass (pairs with |r| small against the columns' means,
    # as the reference's raw-sum gate flags them, fc/_cpu_core.py:173-192)
    rng = np.random.default_rng(3)
    _, st = fcb.stock_corr_f64(rng.standard_normal((200, 30)), None, return_stats=True)
    assert 0 <= st["fallback_count"] < 30 * 29 // 2 // 4
    _, st = fcb.stock_corr_f64(rng.standard_normal((200, 30)) + 50.0, None, return_stats=True)
    assert st["fallback_count"] == 30 * 29 // 2  # every pair cancels in its raw sums


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_rolling_ic_return_ranks_binding(dtype):
    fcb = u.fcb()
    rng = np.random.default_rng(8)
    T, N = 45, 700
    F = np.round(rng.standard_normal((T, N)), 2).astype(dtype)
    R = (0.1 * F + rng.standard_normal((T, N))).astype(dtype)
    F[rng.random((T, N)) < 0.03] = np.nan
    R[rng.random((T, N)) < 0.03] = np.inf
    fm = rng.random((T, N)) < 0.9
    rm = rng.random((T, N)) < 0.95
    fm[4] = False                     # an empty date
    ic, rf, rr = fcb.rolling_ic_f64(F, R, fm, rm, 30, return_ranks=True)
    assert rf.shape == (T, N) and rr.shape == (T, N) and rf.dtype == np.float64
    pv = np.isfinite(F) & np.isfinite(R) & fm & rm
    if dtype == np.float32:
        ref_f = c_oracle.rank_rows(F, pv)
        ref_r = c_oracle.rank_rows(R, pv)
    else:
        # the C oracle ranks float32 keys; float64 rows by the numpy arm's
        # stable argsort of the float64 values
        ref_f, ref_r = np.zeros((T, N)), np.zeros((T, N))
        for t in range(T):
            ref_f[t, pv[t]] = np_arm.ordinal_rank_1d(F[t, pv[t]])
            ref_r[t, pv[t]] = np_arm.ordinal_rank_1d(R[t, pv[t]])
    ref_f = np.where(pv, ref_f, 0.0)
    ref_r = np.where(pv, ref_r, 0.0)
    assert np.array_equal(rf, ref_f)
    assert np.array_equal(rr, ref_r)
    assert np.array_equal(ic, c_oracle.rolling_ic(F, R, fm, rm, 30), equal_nan=True)
    # the ranks do not change the IC
    assert np.array_equal(ic, fcb.rolling_ic_f64(F, R, fm, rm, 30), equal_nan=True)


def test_device_route_two_threads_two_streams():
    t = u._torch
    dev = t.device("cuda:0")
    rng = np.random.default_rng(99)
    T, N = 300, 4000
    cases = []
    for k in range(2):
        F = np.round(rng.standard_normal((T, N)), 2).astype(np.float32)
        R = (0.05 * F + rng.standard_normal((T, N))).astype(np.float32)
        fm = rng.random((T, N)) < 0.9
        cases.append((F, R, fm, c_oracle.rolling_ic(F, R, fm, None, 30)))
    errors = []

    def worker(k):
        F, R, fm, want = cases[k]
        s = t.cuda.Stream(device=dev)
        try:
            with t.cuda.stream(s):
                dF = t.from_numpy(F).to(dev, non_blocking=False)
                dR = t.from_numpy(R).to(dev)
                dm = t.from_numpy(fm).to(dev)
                for _ in range(6):
                    got = fc.rolling_ic(dF, dR, dm, None, min_valid=30)
                    g = got.cpu().numpy()
                    if not np.array_equal(g, want, equal_nan=True):
                        errors.append((k, float(np.nanmax(np.abs(g - want)))))
                        return
        except Exception as e:  # reported by the main thread
            errors.append((k, repr(e)))

    th = [threading.Thread(target=worker, args=(k,)) for k in range(2)]
    for x in th:
        x.start()
    for x in th:
        x.join()
    assert not errors, errors
