"""CPU backend of the fc API (backend='cpu', rolling_ic device='cpu' or no
GPU): numpy implementations with the reference CPU arm's algorithms, guards
and fallback order (fc/_cpu_core.py:15-301), so backend='cpu' returns the
reference's numbers. Only reached when the caller asks for the CPU; the GPU
ops never fall back here.
"""
from __future__ import annotations

import numpy as np

_CANCEL_RATIO = 3.0


def in_corr_domain(x, mask=None) -> bool:
    """max|x| <= 1e150 and min nonzero |x| >= 1e-150 over finite (and
    mask-selected) cells; an empty selection is in the domain."""
    x = np.asarray(x, dtype=np.float64)
    sel = np.isfinite(x)
    if mask is not None:
        sel &= np.asarray(mask, dtype=bool)
    v = x[sel]
    if v.size == 0:
        return True
    a = np.abs(v)
    if a.max() > 1e150:
        return False
    nz = a[a != 0.0]
    return not (nz.size and nz.min() < 1e-150)


def _ranks_stable(values) -> np.ndarray:
    r = np.empty(values.size, dtype=np.float64)
    r[np.argsort(values, kind="stable")] = np.arange(1, values.size + 1, dtype=np.float64)
    return r


def np_cs_rank(X, mask, descending: bool) -> np.ndarray:
    X = np.asarray(X, dtype=np.float32)
    if X.ndim != 2:
        raise ValueError(f"cs_rank input must be (T,N) 2-D, got ndim={X.ndim}")
    if X.shape[1] > 2 ** 24:
        raise ValueError(f"N={X.shape[1]} > 2^24: ranks not exactly representable")
    W = -X if descending else X
    sel = np.isfinite(W)
    if mask is not None:
        sel &= np.asarray(mask, dtype=bool)
    out = np.full(X.shape, np.nan, dtype=np.float32)
    for t in np.flatnonzero(sel.any(axis=1)):
        p = sel[t]
        out[t, p] = _ranks_stable(W[t, p]).astype(np.float32)
    return out


def np_parameter_scan(X, mask) -> list:
    return [np_cs_rank(X, mask if m else None, d) for d in (False, True) for m in (True, False)]


def _kahan_mean(x) -> float:
    s = 0.0
    c = 0.0
    for v in np.asarray(x, dtype=np.float64):
        y = v - c
        t = s + y
        c = (t - s) - y
        s = t
    return s / np.asarray(x).size


def _bias_sensitive(x) -> bool:
    x = np.asarray(x, dtype=np.float64)
    return float(abs(x.mean())) > 1e3 * float(x.std())


def _two_pass_corr(xa, xb) -> float:
    a = np.asarray(xa, dtype=np.float64)
    b = np.asarray(xb, dtype=np.float64)
    if a.size < 2:
        return float("nan")
    if _bias_sensitive(a) or _bias_sensitive(b):
        am, bm = _kahan_mean(a), _kahan_mean(b)
    else:
        am, bm = a.mean(), b.mean()
    ac, bc = a - am, b - bm
    sxx = float((ac * ac).sum())
    syy = float((bc * bc).sum())
    sxy = float((ac * bc).sum())
    if not (sxx > 0.0 and syy > 0.0) or not (np.isfinite(sxx) and np.isfinite(syy)):
        return float("nan")
    return float((sxy / np.sqrt(sxx)) / np.sqrt(syy))


def _gemm_corr(X64, valid):
    """One-pass masked sums (float64 GEMMs), cancellation / |r|>1 fallback to
    the pair two-pass, diagonal from the self two-pass, upper triangle
    mirrored."""
    M = valid.astype(np.float64)
    Xm = np.where(valid, X64, 0.0)
    n = M.T @ M
    sx = Xm.T @ M
    sy = M.T @ Xm
    sxy = Xm.T @ Xm
    Xq = Xm * Xm
    sxx = Xq.T @ M
    syy = M.T @ Xq
    ns = np.maximum(n, 1.0)
    with np.errstate(all="ignore"):
        cov = (sxy - sx * sy / ns) / (ns - 1.0)
        vx = (sxx - sx * sx / ns) / (ns - 1.0)
        vy = (syy - sy * sy / ns) / (ns - 1.0)
        corr = cov / np.sqrt(vx * vy)
        cu = sxy - sx * sy / ns
        vxu = sxx - sx * sx / ns
        vyu = syy - sy * sy / ns
        fix = ((np.abs(sx * sy / ns) > _CANCEL_RATIO * np.abs(cu))
               | (np.abs(sx * sx / ns) > _CANCEL_RATIO * np.abs(vxu))
               | (np.abs(sy * sy / ns) > _CANCEL_RATIO * np.abs(vyu)))
        fix |= np.abs(corr) > 1.0
    corr[n < 2] = np.nan
    fix = fix | fix.T
    for i, j in zip(*np.nonzero(np.triu(fix, 1))):
        o = valid[:, i] & valid[:, j]
        v = _two_pass_corr(X64[o, i], X64[o, j]) if o.sum() >= 2 else float("nan")
        corr[i, j] = corr[j, i] = v
    for i in range(X64.shape[1]):
        o = valid[:, i]
        v = _two_pass_corr(X64[o, i], X64[o, i])
        corr[i, i] = 1.0 if np.isfinite(v) else np.nan
    return np.triu(corr) + np.triu(corr, 1).T


def np_stock_corr(X, mask) -> np.ndarray:
    if not in_corr_domain(X, mask):
        raise ValueError("corr 数值域外：max|x|>1e150 或 min 非零|x|<1e-150")
    X64 = np.asarray(X, dtype=np.float64)
    valid = np.isfinite(X64)
    if mask is not None:
        valid &= np.asarray(mask, dtype=bool)
    return _gemm_corr(X64, valid)


def np_factor_corr(F3, mask) -> np.ndarray:
    F3 = np.asarray(F3)
    if F3.ndim != 3:
        raise ValueError(f"factor_corr input must be (T,N,F) 3-D, got ndim={F3.ndim}")
    T, N, F = F3.shape
    X64 = F3.astype(np.float64).reshape(T * N, F)
    valid = np.isfinite(X64)
    if mask is not None:
        valid &= np.asarray(mask, dtype=bool).reshape(-1, 1)
    if not in_corr_domain(X64, valid):
        raise ValueError("corr 数值域外：max|x|>1e150 或 min 非零|x|<1e-150")
    return _gemm_corr(X64, valid)


def np_rolling_ic(f, r, factor_mask, fwd_mask, min_valid: int) -> np.ndarray:
    f = np.asarray(f, dtype=np.float64)
    r = np.asarray(r, dtype=np.float64)
    ok = np.isfinite(f) & np.isfinite(r)
    if factor_mask is not None:
        ok &= np.asarray(factor_mask, dtype=bool)
    if fwd_mask is not None:
        ok &= np.asarray(fwd_mask, dtype=bool)
    out = np.full(f.shape[0], np.nan)
    for t in range(f.shape[0]):
        o = ok[t]
        if o.sum() < min_valid:
            continue
        fv, rv = f[t][o], r[t][o]
        if np.ptp(fv) == 0 or np.ptp(rv) == 0:
            continue
        out[t] = _two_pass_corr(_ranks_stable(fv), _ranks_stable(rv))
    return out
