"""ctypes view of libfc_hip.so's C ABI (include/fc_hip.h) for the tests."""
from __future__ import annotations

import ctypes
import functools
import pathlib
import re

import numpy as np

REPO = pathlib.Path(__file__).resolve().parent.parent
LIB = REPO / "factor-cuda_amd" / "build" / "libfc_hip.so"
HEADER = REPO / "include" / "fc_hip.h"

vp = ctypes.c_void_p
i32 = ctypes.c_int
i64 = ctypes.c_longlong
dbl = ctypes.c_double


class StockStats(ctypes.Structure):
    _fields_ = [("selected_path", ctypes.c_int), ("fallback_count", ctypes.c_int)]


class Panel(ctypes.Structure):
    """fc_panel (include/fc_hip.h)."""
    _fields_ = [("X", ctypes.c_void_p), ("mask", ctypes.c_void_p), ("R", ctypes.c_void_p),
                ("rmask", ctypes.c_void_p), ("T", ctypes.c_int), ("N", ctypes.c_int),
                ("F", ctypes.c_int), ("min_valid", ctypes.c_int), ("rank_out", ctypes.c_void_p),
                ("ic", ctypes.c_void_p), ("corr", ctypes.c_void_p)]


class MemStats(ctypes.Structure):
    _fields_ = [("live_bytes", ctypes.c_int64), ("peak_bytes", ctypes.c_int64),
                ("n_alloc", ctypes.c_int64), ("n_free", ctypes.c_int64)]


SIGS = {
    "fc_workspace_create": ([vp], i32),
    "fc_workspace_clear": ([vp], None),
    "fc_workspace_destroy": ([vp], None),
    "fc_cs_rank_f32": ([vp, vp, i32, i32, i32, vp, vp], i32),
    "fc_parameter_scan_f32": ([vp, vp, i32, i32, vp, vp, vp, vp, vp, vp], i32),
    "fc_rolling_ic": ([vp, vp, i32, vp, vp, i32, i32, i32, vp, vp, vp, vp], i32),
    "fc_stock_corr": ([vp, i32, vp, i32, i32, i32, vp, vp], i32),
    "fc_factor_corr": ([vp, i32, vp, i32, i32, i32, i32, vp], i32),
    "fc_rank_scan_dev": ([vp, i32, i32, i32, i64, i64, i64, vp, i64, vp, vp, vp, vp, vp], i32),
    "fc_rank_ic_dev": ([vp, i32, i32, i32, i32, i64, i64, i64, vp, i64, vp, vp, i32, vp, vp, vp,
                        vp], i32),
    "fc_fwd_order_dev": ([vp, i32, vp, i32, i32, vp, vp], i32),
    "fc_rank_ic_dev_noalloc": ([vp, i32, i32, i32, i32, i64, i64, i64, vp, i64, vp, vp, i32, vp,
                                vp, vp], i32),
    "fc_rank_ic_noalloc_scratch_bytes": ([], i64),
    "fc_rank_ic_ordered_dev": ([vp, i32, i32, i32, i32, i64, i64, i64, vp, i64, i32, vp, vp, vp,
                                vp], i32),
    "fc_stock_corr_dev": ([vp, vp, i32, i32, i32, vp, vp, vp], i32),
    "fc_factor_corr_dev": ([vp, vp, i32, i32, i32, i32, vp, vp, vp], i32),
    "fc_stock_corr_dev_ex": ([vp, i32, vp, i32, i32, i32, vp, vp, vp, vp], i32),
    "fc_factor_corr_dev_ex": ([vp, i32, vp, i32, i32, i32, i32, vp, vp, vp], i32),
    "fc_batch_create": ([vp, i32, i32, vp], i32),
    "fc_batch_run": ([vp, vp], i32),
    "fc_batch_run_panel": ([vp, i32, vp], i32),
    "fc_batch_tune": ([vp, i32, vp], i32),
    "fc_batch_get_tuning": ([vp, vp, vp], i32),
    "fc_batch_set_tuning": ([vp, i32, i32], i32),
    "fc_batch_destroy": ([vp], None),
    "fc_mem_get_stats": ([vp], None),
    "fc_mem_reset_peak": ([], None),
    "fc_rank_max_n": ([i32], i32),
    "fc_op_device_bytes": ([i32, i32, i32, i32, i32, i32, i32], i64),
    "fc_mem_info": ([vp, vp], i32),
    "fc_version": ([], ctypes.c_char_p),
}


def header_symbols():
    text = HEADER.read_text()
    return sorted(set(re.findall(r"\b(fc_[a-z0-9_]+)\s*\(", text)))


@functools.lru_cache(maxsize=1)
def lib():
    L = ctypes.CDLL(str(LIB))
    for name, (args, res) in SIGS.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = res
    return L


def ptr(a):
    if a is None:
        return None
    if hasattr(a, "data_ptr"):
        return ctypes.c_void_p(a.data_ptr())
    return a.ctypes.data_as(ctypes.c_void_p)


class Workspace:
    def __init__(self):
        self.h = ctypes.c_void_p()
        rc = lib().fc_workspace_create(ctypes.byref(self.h))
        assert rc == 0

    def __del__(self):
        try:
            lib().fc_workspace_destroy(self.h)
        except Exception:
            pass


def cs_rank(X, mask=None, descending=False, ws=None):
    X = np.ascontiguousarray(X, dtype=np.float32)
    T, N = X.shape
    m = None if mask is None else np.ascontiguousarray(mask, dtype=np.uint8)
    out = np.empty((T, N), dtype=np.float32)
    rc = lib().fc_cs_rank_f32(ptr(X), ptr(m), T, N, int(descending), ptr(out),
                              ws.h if ws else None)
    return rc, out


def parameter_scan(X, mask, active=None, ws=None):
    X = np.ascontiguousarray(X, dtype=np.float32)
    T, N = X.shape
    m = None if mask is None else np.ascontiguousarray(mask, dtype=np.uint8)
    outs = [np.empty((T, N), dtype=np.float32) for _ in range(4)]
    arr = (ctypes.c_void_p * 4)(*[o.ctypes.data for o in outs])
    st = (ctypes.c_int * 4)()
    tms = (ctypes.c_double * 4)()
    tg = (ctypes.c_float * 4)()
    act = None if active is None else (ctypes.c_int * 4)(*active)
    rc = lib().fc_parameter_scan_f32(ptr(X), ptr(m), T, N, arr, st, tms, tg, act,
                                     ws.h if ws else None)
    return rc, outs, list(st), list(tms), list(tg)


def rolling_ic(F, R, fmask=None, rmask=None, min_valid=30, ranks=False, ws=None):
    vb = 8 if F.dtype == np.float64 else 4
    dt = np.float64 if vb == 8 else np.float32
    F = np.ascontiguousarray(F, dtype=dt)
    R = np.ascontiguousarray(R, dtype=dt)
    T, N = F.shape
    fm = None if fmask is None else np.ascontiguousarray(fmask, dtype=np.uint8)
    rm = None if rmask is None else np.ascontiguousarray(rmask, dtype=np.uint8)
    ic = np.empty(T, dtype=np.float64)
    rf = np.empty((T, N)) if ranks else None
    rr = np.empty((T, N)) if ranks else None
    rc = lib().fc_rolling_ic(ptr(F), ptr(R), vb, ptr(fm), ptr(rm), T, N, int(min_valid), ptr(ic),
                             ptr(rf), ptr(rr), ws.h if ws else None)
    return (rc, ic, rf, rr) if ranks else (rc, ic)


def stock_corr(X, mask=None, precision=0):
    vb = 8 if X.dtype == np.float64 else 4
    X = np.ascontiguousarray(X)
    T, N = X.shape
    m = None if mask is None else np.ascontiguousarray(mask, dtype=np.uint8)
    out = np.empty((N, N), dtype=np.float64)
    st = StockStats()
    rc = lib().fc_stock_corr(ptr(X), vb, ptr(m), T, N, int(precision), ptr(out), ctypes.byref(st))
    return rc, out, (st.selected_path, st.fallback_count)


def factor_corr(F3, mask=None, precision=0):
    vb = 8 if F3.dtype == np.float64 else 4
    F3 = np.ascontiguousarray(F3)
    T, N, F = F3.shape
    m = None if mask is None else np.ascontiguousarray(mask, dtype=np.uint8)
    out = np.empty((F, F), dtype=np.float64)
    rc = lib().fc_factor_corr(ptr(F3), vb, ptr(m), T, N, F, int(precision), ptr(out))
    return rc, out


def mem_stats():
    s = MemStats()
    lib().fc_mem_get_stats(ctypes.byref(s))
    return s


OP_CS_RANK, OP_PARAMETER_SCAN, OP_ROLLING_IC, OP_STOCK_CORR, OP_FACTOR_CORR = range(5)
PLAN_MASK, PLAN_RMASK, PLAN_RANKS = 1, 2, 4


def op_bytes(op, T, N, F=1, vb=4, precision=0, flags=0):
    return lib().fc_op_device_bytes(op, T, N, F, vb, precision, flags)
