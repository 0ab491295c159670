"""ctypes binding of the C oracle (test infrastructure only)."""
from __future__ import annotations

import ctypes
import functools
import pathlib

import numpy as np

_LIB = pathlib.Path(__file__).resolve().parent / "_build" / "liboracle_c.so"


@functools.lru_cache(maxsize=1)
def lib():
    if not _LIB.exists():
        from . import build_c
        build_c.build()
    L = ctypes.CDLL(str(_LIB))
    vp = ctypes.c_void_p
    L.oracle_rank_rows.argtypes = [vp, vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, vp]
    L.oracle_rank_rows.restype = None
    L.oracle_rolling_ic.argtypes = [vp, vp, ctypes.c_int, vp, vp, ctypes.c_int, ctypes.c_int,
                                    ctypes.c_int, vp]
    L.oracle_rolling_ic.restype = None
    return L


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def rank_rows(X, mask=None, descending=False):
    X = np.ascontiguousarray(X, dtype=np.float32)
    T, N = X.shape
    m = None if mask is None else np.ascontiguousarray(mask, dtype=np.uint8)
    out = np.empty((T, N), dtype=np.float32)
    lib().oracle_rank_rows(_p(X), _p(m), T, N, int(descending), _p(out))
    return out


def rolling_ic(F, R, fmask=None, rmask=None, min_valid=30):
    is64 = F.dtype == np.float64
    dt = np.float64 if is64 else np.float32
    F = np.ascontiguousarray(F, dtype=dt)
    R = np.ascontiguousarray(R, dtype=dt)
    T, N = F.shape
    fm = None if fmask is None else np.ascontiguousarray(fmask, dtype=np.uint8)
    rm = None if rmask is None else np.ascontiguousarray(rmask, dtype=np.uint8)
    out = np.empty(T, dtype=np.float64)
    lib().oracle_rolling_ic(_p(F), _p(R), int(is64), _p(fm), _p(rm), T, N, int(min_valid), _p(out))
    return out
