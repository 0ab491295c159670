"""Restatement of the reference numpy CPU arm (fc/_cpu_core.py:15-301).

Test infrastructure (see oracle/__init__.py): the checker for the GPU path and
the CPU baseline timed by bench.py. Same numpy algorithms, guards and
fallback order as the reference arm; each function cites the lines it
restates.
"""
from __future__ import annotations

import numpy as np

CANCEL_RATIO = 3.0  # fc/_cpu_core.py:171


def in_corr_domain(x, mask=None) -> bool:
    """fc/_cpu_core.py:15-41: max|x| <= 1e150 and min nonzero |x| >= 1e-150
    over the finite (and mask-selected) cells; empty -> True."""
    x = np.asarray(x, dtype=np.float64)
    ok = np.isfinite(x)
    if mask is not None:
        ok &= np.asarray(mask, dtype=bool)
    v = x[ok]
    if v.size == 0:
        return True
    if np.max(np.abs(v)) > 1e150:
        return False
    nz = v[v != 0.0]
    if nz.size and np.min(np.abs(nz)) < 1e-150:
        return False
    return True


def ordinal_rank_1d(x) -> np.ndarray:
    """fc/_cpu_core.py:43-55: stable ordinal ranks 1..K of the finite entries
    (input precision kept), NaN elsewhere."""
    x = np.asarray(x)
    out = np.full(x.shape, np.nan, dtype=np.float64)
    fin = np.isfinite(x)
    if fin.any():
        order = np.argsort(x[fin], kind="stable")
        r = np.empty(order.size, dtype=np.float64)
        r[order] = np.arange(1, order.size + 1, dtype=np.float64)
        out[fin] = r
    return out


def np_cs_rank(X, mask, descending: bool) -> np.ndarray:
    """fc/_cpu_core.py:57-78: per row, stable argsort of the float32 values
    (negated for descending) over finite & mask cells; quiet NaN elsewhere."""
    X = np.asarray(X, dtype=np.float32)
    if X.ndim != 2:
        raise ValueError(f"cs_rank input must be (T,N) 2-D, got ndim={X.ndim}")
    if X.shape[1] > 2 ** 24:
        raise ValueError("N > 2^24: ranks not exactly representable")
    W = -X if descending else X
    out = np.full(X.shape, np.nan, dtype=np.float32)
    part = np.isfinite(W)
    if mask is not None:
        part &= np.asarray(mask, dtype=bool)
    for t in range(X.shape[0]):
        p = part[t]
        if not p.any():
            continue
        order = np.argsort(W[t][p], kind="stable")
        r = np.empty(order.size, dtype=np.float32)
        r[order] = np.arange(1, order.size + 1, dtype=np.float32)
        out[t][p] = r
    return out


def np_parameter_scan(X, mask) -> list:
    """fc/_cpu_core.py:295-301: groups in binding order asc-masked,
    asc-unmasked, desc-masked, desc-unmasked."""
    return [np_cs_rank(X, mask if use_mask else None, desc)
            for desc in (False, True) for use_mask in (True, False)]


def kahan_mean(x) -> float:
    """fc/_cpu_core.py:80-96: serial Kahan-compensated mean."""
    s = 0.0
    c = 0.0
    for v in np.asarray(x, dtype=np.float64):
        y = v - c
        t = s + y
        c = (t - s) - y
        s = t
    return s / np.asarray(x).size


def bias_sensitive(x) -> bool:
    """fc/_cpu_core.py:99-108: |mean| > 1e3 * std (ddof=0)."""
    x = np.asarray(x, dtype=np.float64)
    return float(abs(x.mean())) > 1e3 * float(x.std())


def two_pass_corr(a, b) -> float:
    """fc/_cpu_core.py:111-137: two-pass centred Pearson (ddof=1 cancels),
    Kahan mean on bias-sensitive inputs, safe order (sxy/sqrt(sxx))/sqrt(syy),
    NaN on non-positive or non-finite sums of squares."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    if a.size < 2:
        return float("nan")
    if bias_sensitive(a) or bias_sensitive(b):
        am, bm = kahan_mean(a), kahan_mean(b)
    else:
        am, bm = a.mean(), b.mean()
    ac = a - am
    bc = b - bm
    sxx = float((ac * ac).sum())
    syy = float((bc * bc).sum())
    sxy = float((ac * bc).sum())
    if not (sxx > 0.0 and syy > 0.0) or not (np.isfinite(sxx) and np.isfinite(syy)):
        return float("nan")
    return float((sxy / np.sqrt(sxx)) / np.sqrt(syy))


def masked_gemm_stats(X, mask):
    """fc/_cpu_core.py:139-169: one-pass masked GEMM sums in float64."""
    X64 = X.astype(np.float64, copy=False)
    valid = np.isfinite(X64)
    if mask is not None:
        valid &= np.asarray(mask, dtype=bool)
    M = valid.astype(np.float64)
    Xm = np.where(valid, X64, 0.0)
    n = M.T @ M
    sumx = Xm.T @ M
    sumy = M.T @ Xm
    sumxy = Xm.T @ Xm
    Xsq = Xm * Xm
    sumx2 = Xsq.T @ M
    sumy2 = M.T @ Xsq
    n_safe = np.maximum(n, 1.0)
    den = n_safe - 1.0
    with np.errstate(all="ignore"):
        cov = (sumxy - sumx * sumy / n_safe) / den
        varx = (sumx2 - sumx * sumx / n_safe) / den
        vary = (sumy2 - sumy * sumy / n_safe) / den
        corr = cov / np.sqrt(varx * vary)
    corr[n < 2] = np.nan
    return corr, valid, n, sumx, sumy, sumxy, sumx2, sumy2


def gemm_cancel_mask(sumx, sumy, sumxy, sumx2, sumy2, n):
    """fc/_cpu_core.py:173-192: cells whose one-pass sums cancel by more than
    the ratio threshold."""
    n_safe = np.maximum(n, 1.0)
    cov_u = sumxy - sumx * sumy / n_safe
    varx_u = sumx2 - sumx * sumx / n_safe
    vary_u = sumy2 - sumy * sumy / n_safe
    with np.errstate(all="ignore"):
        fix = ((np.abs(sumx * sumy / n_safe) > CANCEL_RATIO * np.abs(cov_u))
               | (np.abs(sumx * sumx / n_safe) > CANCEL_RATIO * np.abs(varx_u))
               | (np.abs(sumy * sumy / n_safe) > CANCEL_RATIO * np.abs(vary_u)))
    return fix | fix.T


def _finish(X64, valid, stats):
    """Shared tail of fc/_cpu_core.py:206-230 and :250-269: cancellation and
    |r|>1 fallback to the pair-specific two-pass, diagonal from the computed
    self-correlation, upper triangle mirrored."""
    corr, _, n, sumx, sumy, sumxy, sumx2, sumy2 = stats
    fix = gemm_cancel_mask(sumx, sumy, sumxy, sumx2, sumy2, n)
    with np.errstate(invalid="ignore"):
        fix |= np.abs(corr) > 1.0
    fix = fix | fix.T
    for i, j in zip(*np.where(fix)):
        if i >= j:
            continue
        o = valid[:, i] & valid[:, j]
        v = np.nan if o.sum() < 2 else two_pass_corr(X64[o, i], X64[o, j])
        corr[i, j] = corr[j, i] = v
    for i in range(X64.shape[1]):
        o = valid[:, i]
        v = two_pass_corr(X64[o, i], X64[o, i])
        corr[i, i] = 1.0 if np.isfinite(v) else np.nan
    return np.triu(corr) + np.triu(corr, 1).T


def np_stock_corr(X, mask) -> np.ndarray:
    """fc/_cpu_core.py:194-230."""
    if not in_corr_domain(X, mask):
        raise ValueError("corr value domain violated (max|x|>1e150 or min nonzero |x|<1e-150)")
    X = np.asarray(X, dtype=np.float64)
    stats = masked_gemm_stats(X, mask)
    return _finish(X, stats[1], stats)


def np_factor_corr(F3, mask) -> np.ndarray:
    """fc/_cpu_core.py:232-269: (T,N,F) pooled to (T*N, F)."""
    F3 = np.asarray(F3)
    if F3.ndim != 3:
        raise ValueError(f"factor_corr input must be (T,N,F) 3-D, got ndim={F3.ndim}")
    T, N, F = F3.shape
    X64 = F3.astype(np.float64).reshape(T * N, F)
    valid = np.isfinite(X64)
    if mask is not None:
        valid &= np.asarray(mask, dtype=bool).reshape(-1, 1)
    if not in_corr_domain(X64, valid):
        raise ValueError("corr value domain violated (max|x|>1e150 or min nonzero |x|<1e-150)")
    stats = masked_gemm_stats(X64, valid)
    return _finish(X64, stats[1], stats)


def np_rolling_ic(f, r, factor_mask, fwd_mask, min_valid: int) -> np.ndarray:
    """fc/_cpu_core.py:271-293: per date, Spearman on the pairwise-valid set
    (finite both sides, both masks), NaN below min_valid or on a constant
    side."""
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
        if np.ptp(f[t][o]) == 0 or np.ptp(r[t][o]) == 0:
            continue
        out[t] = two_pass_corr(ordinal_rank_1d(f[t][o]), ordinal_rank_1d(r[t][o]))
    return out
