"""GPU tests aimed at the bucketed rank of float rows in rank + IC
(fc_rank_ic_ordered_dev with float32 values): rows chosen to stress the
monotone bucket map and the switch to the radix sort for crowded buckets,
each plane's masked ascending ranks bitwise and its IC exactly against the C
oracle (reference semantics: src/cross_sectional_rank.cu:35-72,
src/rolling_ic.cu:183-540).

Row patterns, one per date: rounded normals (ties), values one ulp apart,
subnormal-only spreads (an infinite map scale), +-FLT_MAX next to ordinary
values, +-0 mixes, a constant row, two values only, lognormal and Cauchy
tails, rows of a handful of valid items, and a value repeated 46 and 49 times
(around the bucket size of 48 above which a row goes to the radix sort)."""
import ctypes

import numpy as np
import pytest

import capi
from oracle import c_oracle

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

vp = ctypes.c_void_p
F32 = np.float32


def _row(kind, rng, N):
    if kind == "round3":
        return np.round(rng.standard_normal(N), 3).astype(F32)
    if kind == "ulps":
        return (F32(1.0) + np.arange(N, dtype=F32)[rng.permutation(N)] * np.finfo(F32).eps
                ).astype(F32)
    if kind == "subnormal":
        return (rng.integers(-6, 7, N).astype(F32) * np.finfo(F32).smallest_subnormal).astype(F32)
    if kind == "extremes":
        x = rng.standard_normal(N).astype(F32)
        x[rng.integers(0, N, 3)] = np.finfo(F32).max
        x[rng.integers(0, N, 3)] = -np.finfo(F32).max
        return x
    if kind == "zeros":
        x = rng.standard_normal(N).astype(F32)
        z = rng.random(N) < 0.3
        x[z] = np.where(rng.random(z.sum()) < 0.5, F32(0.0), F32(-0.0))
        return x
    if kind == "constant":
        return np.full(N, F32(2.5))
    if kind == "two":
        return np.where(rng.random(N) < 0.5, F32(-1.0), F32(3.0)).astype(F32)
    if kind == "lognormal":
        return rng.lognormal(0.0, 2.0, N).astype(F32)
    if kind == "cauchy":
        return rng.standard_cauchy(N).astype(F32)
    if kind in ("eq46", "eq49"):
        # a spread of distinct values, one of them repeated 46 / 49 times:
        # with a neighbour or two in its bucket, around the bucket size at
        # which a row switches to the radix sort
        x = np.linspace(-1.0, 1.0, N).astype(F32)[rng.permutation(N)]
        k = 46 if kind == "eq46" else 49
        x[1 + rng.choice(N - 1, k, replace=False)] = F32(0.123)
        return x
    raise ValueError(kind)


KINDS = ["round3", "ulps", "subnormal", "extremes", "zeros", "constant", "two", "lognormal",
         "cauchy", "eq46", "eq49", "sparse"]


def _panel(rng, N, F):
    T = len(KINDS)
    X = np.empty((F, T, N), F32)
    M = rng.random((T, N)) < 0.85
    for t, kind in enumerate(KINDS):
        for f in range(F):
            X[f, t] = _row("round3" if kind == "sparse" else kind, rng, N)
        if kind == "sparse":
            M[t] = False
            M[t, rng.choice(N, 5, replace=False)] = True
        if kind in ("eq46", "eq49"):
            M[t] = True
    X[:, :, 0] = np.nan  # a non-finite cell in every row
    R = (0.1 * np.nan_to_num(X[0], posinf=0, neginf=0) + rng.standard_normal((T, N))).astype(F32)
    R[3] = np.round(R[3])  # return ties
    RM = rng.random((T, N)) < 0.95
    return X, M, R, RM


@pytest.mark.parametrize("N", [300, 5000, 12288])
def test_bucket_rows_match_oracle(N):
    rng = np.random.default_rng(N)
    F = 2
    X, M, R, RM = _panel(rng, N, F)
    T = X.shape[1]
    dev = "cuda:0"
    dX = torch.from_numpy(X).to(dev)
    dM = torch.from_numpy(M.view(np.uint8)).to(dev)
    dR = torch.from_numpy(R).to(dev)
    dRM = torch.from_numpy(RM.view(np.uint8)).to(dev)
    ic = torch.full((F, T), -7.0, dtype=torch.float64, device=dev)
    rk = torch.full((F, T, N), -7.0, dtype=torch.float32, device=dev)
    L = capi.lib()
    ws = capi.Workspace()
    assert L.fc_fwd_order_dev(vp(dR.data_ptr()), 4, vp(dRM.data_ptr()), T, N, ws.h, None) == 0
    for rank_out in (True, False):
        rc = L.fc_rank_ic_ordered_dev(vp(dX.data_ptr()), 4, F, T, N, T * N, N, 1,
                                      vp(dM.data_ptr()), 0, 2, vp(ic.data_ptr()),
                                      vp(rk.data_ptr()) if rank_out else None, ws.h, None)
        assert rc == 0, rc
        torch.cuda.synchronize()
        got_ic, got_rk = ic.cpu().numpy(), rk.cpu().numpy()
        for f in range(F):
            want = c_oracle.rolling_ic(X[f], R, M, RM, 2)
            assert np.array_equal(got_ic[f], want, equal_nan=True), (f, KINDS)
            if rank_out:
                want_rk = c_oracle.rank_rows(X[f], M)
                bad = np.nonzero(~np.all(got_rk[f].view(np.uint32) == want_rk.view(np.uint32),
                                         axis=1))[0]
                assert bad.size == 0, [KINDS[t] for t in bad]
