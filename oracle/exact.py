"""Independent definitions that pin oracle.np_arm and the GPU path.

Test infrastructure (see oracle/__init__.py). These follow the contract's
definitions, not the numpy arm's algorithms, so a mistake shared by the arm
and the GPU path (a dropped term, a transposed operand, a wrong tie order)
fails here.
"""
from __future__ import annotations

import math
from fractions import Fraction

import numpy as np


def f32_key(v: float) -> int:
    """Canonical ordinal key of a float32 (tests/fixtures/corr_math_v1.py:189-195):
    +0/-0 fold, sign-flip transform, non-finite -> 0xFFFFFFFF."""
    v = np.float32(v)
    if not np.isfinite(v):
        return 0xFFFFFFFF
    b = 0 if v == 0.0 else int(np.array(v, dtype=np.float32).view(np.uint32))
    return (~b) & 0xFFFFFFFF if b & 0x80000000 else b | 0x80000000


def brute_rank_row(x, valid, descending=False) -> np.ndarray:
    """Contract definition (CLAUDE.md:63), O(N^2):
    asc[i] = 1 + #{j valid: y_j < y_i or (y_j == y_i and j < i)} with
    y = -x for descending; NaN for invalid cells."""
    x = np.asarray(x, dtype=np.float32)
    y = -x if descending else x
    n = x.size
    out = np.full(n, np.nan, dtype=np.float32)
    idx = [i for i in range(n) if valid[i]]
    for i in idx:
        c = 1
        for j in idx:
            if y[j] < y[i] or (y[j] == y[i] and j < i):
                c += 1
        out[i] = c
    return out


def stable_rank_row(x, valid, descending=False) -> np.ndarray:
    """Same definition through a stable argsort of the canonical keys."""
    x = np.asarray(x, dtype=np.float32)
    y = -x if descending else x
    out = np.full(x.size, np.nan, dtype=np.float32)
    idx = np.flatnonzero(valid)
    if idx.size:
        keys = np.array([f32_key(v) for v in y[idx]], dtype=np.uint64)
        order = np.argsort(keys, kind="stable")
        r = np.empty(idx.size, dtype=np.float32)
        r[order] = np.arange(1, idx.size + 1, dtype=np.float32)
        out[idx] = r
    return out


def spearman_exact(f, r, valid, min_valid) -> float:
    """Spearman IC of ordinal ranks with exact rational arithmetic, rounded
    once at the end: IC = 3(4S - m(m+1)^2) / (m(m^2-1)), S = sum rf*rr.
    NaN below min_valid or when either side is constant on the valid set."""
    f = np.asarray(f, dtype=np.float64)[valid]
    r = np.asarray(r, dtype=np.float64)[valid]
    m = f.size
    if m < min_valid or m < 2:
        return float("nan")
    if np.ptp(f) == 0 or np.ptp(r) == 0:
        return float("nan")
    rf = np.empty(m, dtype=np.int64)
    rf[np.argsort(f, kind="stable")] = np.arange(1, m + 1)
    rr = np.empty(m, dtype=np.int64)
    rr[np.argsort(r, kind="stable")] = np.arange(1, m + 1)
    S = int(np.sum(rf * rr))
    return float(Fraction(3 * (4 * S - m * (m + 1) ** 2), m * (m * m - 1)))


def spearman_corrcoef(f, r, valid, min_valid) -> float:
    """np.corrcoef of the ordinal ranks (tests/test_adapter_v1.py:1142-1167
    style independent oracle)."""
    f = np.asarray(f, dtype=np.float64)[valid]
    r = np.asarray(r, dtype=np.float64)[valid]
    if f.size < min_valid or f.size < 2 or np.ptp(f) == 0 or np.ptp(r) == 0:
        return float("nan")
    rf = np.empty(f.size)
    rf[np.argsort(f, kind="stable")] = np.arange(1, f.size + 1)
    rr = np.empty(r.size)
    rr[np.argsort(r, kind="stable")] = np.arange(1, r.size + 1)
    return float(np.corrcoef(np.stack([rf, rr]))[0, 1])


def pair_corrcoef(a, b, valid) -> float:
    """Per-pair oracle: np.corrcoef on the pairwise-valid subset
    (tests/fixtures/corr_oracle_v1.py:55-88); n < 2 or exact-constant side
    -> NaN."""
    a = np.asarray(a, dtype=np.float64)[valid]
    b = np.asarray(b, dtype=np.float64)[valid]
    if a.size < 2 or np.ptp(a) == 0 or np.ptp(b) == 0:
        return float("nan")
    with np.errstate(all="ignore"):
        v = float(np.corrcoef(np.stack([a, b]))[0, 1])
    return v


def pearson_fraction(a, b) -> float:
    """Mathematically exact Pearson of the given float64 values (Fraction
    sums, one final rounding per sqrt)."""
    A = [Fraction(float(v)) for v in a]
    B = [Fraction(float(v)) for v in b]
    n = len(A)
    if n < 2:
        return float("nan")
    ma = sum(A) / n
    mb = sum(B) / n
    sxx = sum((x - ma) ** 2 for x in A)
    syy = sum((y - mb) ** 2 for y in B)
    sxy = sum((x - ma) * (y - mb) for x, y in zip(A, B))
    if sxx == 0 or syy == 0:
        return float("nan")
    return float(sxy) / math.sqrt(float(sxx)) / math.sqrt(float(syy))


def kahan_sum(x) -> float:
    s = 0.0
    c = 0.0
    for v in np.asarray(x, dtype=np.float64):
        y = v - c
        t = s + y
        c = (t - s) - y
        s = t
    return s


def serial_kahan_corr(a, b) -> float:
    """HG-2 high-precision reference (tests/test_adapter_v1.py:96-124):
    serial Kahan means and serial Kahan centred sums."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    n = a.size
    if n < 2:
        return float("nan")
    am = kahan_sum(a) / n
    bm = kahan_sum(b) / n
    ac = a - am
    bc = b - bm
    sxx = kahan_sum(ac * ac)
    syy = kahan_sum(bc * bc)
    sxy = kahan_sum(ac * bc)
    if sxx <= 0 or syy <= 0 or not (np.isfinite(sxx) and np.isfinite(syy)):
        return float("nan")
    return float((sxy / np.sqrt(sxx)) / np.sqrt(syy))


def corr_matrix_pairwise(X, mask=None) -> np.ndarray:
    """Full (N,N) matrix from the per-pair oracle; diagonal 1.0 iff the
    self-correlation is finite; upper triangle mirrored."""
    X = np.asarray(X, dtype=np.float64)
    valid = np.isfinite(X)
    if mask is not None:
        valid &= np.asarray(mask, dtype=bool)
    N = X.shape[1]
    out = np.full((N, N), np.nan)
    for i in range(N):
        for j in range(i, N):
            o = valid[:, i] & valid[:, j]
            v = pair_corrcoef(X[:, i], X[:, j], o)
            if i == j:
                v = 1.0 if np.isfinite(v) else np.nan
            out[i, j] = out[j, i] = v
    return out
