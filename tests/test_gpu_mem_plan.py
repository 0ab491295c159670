"""Allocator accounting against the per-op plan (fc_op_device_bytes), on the
GPU: one host-driver call in a new workspace raises the tracker's
high-water mark by exactly the planned bytes and leaves nothing live (the
reference's formula-vs-HWM calibration, benchmarks/results/poc3_calibration_v1.md:8-13,
with delta 0 as the pass bar); a second call in a cached workspace
allocates nothing; a plan beyond free HBM is refused before any allocation
(in place of the fixed budget of benchmarks/memory_budget_v1.py:47)."""
import ctypes

import numpy as np
import pytest

import capi

pytestmark = pytest.mark.gpu


def _hwm_of(call):
    L = capi.lib()
    s0 = capi.mem_stats()
    L.fc_mem_reset_peak()
    rc = call()
    s1 = capi.mem_stats()
    assert rc == 0, rc
    assert s1.live_bytes == s0.live_bytes, "the call left device memory live"
    return s1.peak_bytes - s0.live_bytes


def _panel(rng, T, N, dtype=np.float32):
    x = rng.standard_normal((T, N)).astype(dtype)
    x[rng.random((T, N)) < 0.02] = np.nan
    return x


@pytest.mark.parametrize("T,N", [(37, 1001), (250, 5000)])
@pytest.mark.parametrize("masked", [False, True])
def test_cs_rank_hwm_equals_plan(rng, T, N, masked):
    x = _panel(rng, T, N)
    m = rng.random((T, N)) < 0.9 if masked else None
    hwm = _hwm_of(lambda: capi.cs_rank(x, m)[0])
    assert hwm == capi.op_bytes(capi.OP_CS_RANK, T, N, flags=capi.PLAN_MASK if masked else 0)


@pytest.mark.parametrize("active", [None, [0, 1, 0, 1], [1, 0, 0, 0]])
def test_parameter_scan_hwm_equals_plan(rng, active):
    T, N = 61, 2300
    x = _panel(rng, T, N)
    m = rng.random((T, N)) < 0.9
    hwm = _hwm_of(lambda: capi.parameter_scan(x, m, active=active)[0])
    groups = 15 if active is None else sum(1 << g for g in range(4) if active[g])
    assert hwm == capi.op_bytes(capi.OP_PARAMETER_SCAN, T, N, flags=groups << 8)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("fm,rm,ranks", [(False, False, False), (True, True, False),
                                         (True, False, True)])
def test_rolling_ic_hwm_equals_plan(rng, dtype, fm, rm, ranks):
    T, N = 40, 3100
    f = _panel(rng, T, N, dtype)
    r = _panel(rng, T, N, dtype)
    fmask = rng.random((T, N)) < 0.9 if fm else None
    rmask = rng.random((T, N)) < 0.9 if rm else None
    hwm = _hwm_of(lambda: capi.rolling_ic(f, r, fmask, rmask, 30, ranks=ranks)[0])
    flags = (capi.PLAN_MASK if fm else 0) | (capi.PLAN_RMASK if rm else 0) | \
        (capi.PLAN_RANKS if ranks else 0)
    assert hwm == capi.op_bytes(capi.OP_ROLLING_IC, T, N, vb=np.dtype(dtype).itemsize, flags=flags)


@pytest.mark.parametrize("precision", [0, 1, 2])
@pytest.mark.parametrize("N,dtype,masked", [(300, np.float32, True), (1000, np.float64, False)])
def test_stock_corr_hwm_equals_plan(rng, precision, N, dtype, masked):
    T = 130
    x = _panel(rng, T, N, dtype)
    m = rng.random((T, N)) < 0.9 if masked else None
    hwm = _hwm_of(lambda: capi.stock_corr(x, m, precision)[0])
    assert hwm == capi.op_bytes(capi.OP_STOCK_CORR, T, N, vb=np.dtype(dtype).itemsize,
                                precision=precision, flags=capi.PLAN_MASK if masked else 0)


@pytest.mark.parametrize("precision", [0, 1, 2])
@pytest.mark.parametrize("F,masked", [(12, True), (128, False)])
def test_factor_corr_hwm_equals_plan(rng, precision, F, masked):
    T, N = 50, 400
    x = rng.standard_normal((T, N, F)).astype(np.float32)
    m = rng.random((T, N)) < 0.9 if masked else None
    hwm = _hwm_of(lambda: capi.factor_corr(x, m, precision)[0])
    assert hwm == capi.op_bytes(capi.OP_FACTOR_CORR, T, N, F=F, precision=precision,
                                flags=capi.PLAN_MASK if masked else 0)


def test_cached_workspace_allocates_once(rng):
    T, N = 45, 2000
    x = _panel(rng, T, N)
    ws = capi.Workspace()
    L = capi.lib()
    s0 = capi.mem_stats()
    assert capi.cs_rank(x, None, ws=ws)[0] == 0
    s1 = capi.mem_stats()
    assert s1.live_bytes - s0.live_bytes == capi.op_bytes(capi.OP_CS_RANK, T, N)
    assert capi.cs_rank(x, None, ws=ws)[0] == 0
    s2 = capi.mem_stats()
    assert s2.n_alloc == s1.n_alloc and s2.live_bytes == s1.live_bytes
    L.fc_workspace_clear(ws.h)
    assert capi.mem_stats().live_bytes == s0.live_bytes


def test_plan_beyond_free_hbm_is_refused_before_allocating():
    L = capi.lib()
    free_b, total_b = ctypes.c_size_t(), ctypes.c_size_t()
    assert L.fc_mem_info(ctypes.byref(free_b), ctypes.byref(total_b)) == 0
    assert 0 < free_b.value <= total_b.value
    # T*N at the cap with F = 128: about 1.1 TB of input alone. The driver
    # refuses it from the plan, before it reads the (tiny) host buffer.
    T = N = 46340
    need = capi.op_bytes(capi.OP_FACTOR_CORR, T, N, F=128)
    assert need > free_b.value
    x = np.zeros(4, np.float32)
    out = np.zeros((128, 128))
    s0 = capi.mem_stats()
    rc = L.fc_factor_corr(capi.ptr(x), 4, None, T, N, 128, 0, capi.ptr(out))
    assert rc == 2  # hipErrorOutOfMemory
    s1 = capi.mem_stats()
    assert s1.n_alloc == s0.n_alloc and s1.live_bytes == s0.live_bytes
