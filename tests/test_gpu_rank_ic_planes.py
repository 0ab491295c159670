"""GPU tests of the batched F-plane rank + Spearman IC path through the C ABI
(the benchmark's hot path): fc_fwd_order_dev + fc_rank_ic_ordered_dev over F
factor planes in one launch, every plane's IC exactly against the C oracle's
integer IC and its masked ascending ranks bitwise against the oracle's
ranks (reference semantics: src/rolling_ic.cu:183-540,
src/cross_sectional_rank.cu:35-72), for plane-major (F,T,N) and
factor-fastest (T,N,F) layouts, f32 and f64; and the one-call form
fc_rank_ic_dev bitwise equal to the two-call form, also across repeated
calls, shape changes and a cleared workspace."""
import ctypes

import numpy as np
import pytest

import capi
from oracle import c_oracle

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

vp = ctypes.c_void_p


def _inputs(rng, T, N, F, dtype):
    C = np.atleast_2d(np.corrcoef(rng.standard_normal((F, 3 * F))))
    L = np.linalg.cholesky(C)
    X = np.round(rng.standard_normal((T, N, F)) @ L.T, 3).astype(dtype)
    X[rng.random((T, N, F)) < 0.01] = np.nan
    p = rng.uniform(0.6, 1.0, size=T)
    M = rng.random((T, N)) < p[:, None]
    R = (0.05 * X[..., 0] + rng.standard_normal((T, N))).astype(dtype)
    R[rng.random((T, N)) < 0.01] = np.nan
    RM = rng.random((T, N)) < 0.95
    return X, M, R, RM


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0")


def _strides(layout, T, N, F):
    # (plane, row, element) strides in elements
    return (T * N, N, 1) if layout == "FTN" else (1, N * F, F)


def _run(fn, X, M, R, RM, layout, min_valid=30, ranks=True):
    T, N, F = X.shape
    vb = X.dtype.itemsize
    dX = _dev(np.moveaxis(X, 2, 0) if layout == "FTN" else X)
    dM, dR, dRM = _dev(M.view(np.uint8)), _dev(R), _dev(RM.view(np.uint8))
    ic = torch.full((F, T), -7.0, dtype=torch.float64, device="cuda:0")
    rk = torch.full((F, T, N), -7.0, dtype=torch.float32, device="cuda:0") if ranks else None
    ps, rs, es = _strides(layout, T, N, F)
    ws = capi.Workspace()
    L = capi.lib()
    rkp = vp(rk.data_ptr()) if ranks else None
    if fn == "ordered":
        assert L.fc_fwd_order_dev(vp(dR.data_ptr()), vb, vp(dRM.data_ptr()), T, N, ws.h, None) == 0
        rc = L.fc_rank_ic_ordered_dev(vp(dX.data_ptr()), vb, F, T, N, ps, rs, es, vp(dM.data_ptr()),
                                      0, min_valid, vp(ic.data_ptr()), rkp, ws.h, None)
    else:
        rc = L.fc_rank_ic_dev(vp(dX.data_ptr()), vb, F, T, N, ps, rs, es, vp(dM.data_ptr()), 0,
                              vp(dR.data_ptr()), vp(dRM.data_ptr()), min_valid, vp(ic.data_ptr()),
                              rkp, ws.h, None)
    assert rc == 0, rc
    torch.cuda.synchronize()
    return ic.cpu().numpy(), (rk.cpu().numpy() if ranks else None)


@pytest.mark.parametrize("layout", ["FTN", "TNF"])
@pytest.mark.parametrize("T,N,F", [(40, 5000, 3), (70, 1001, 4), (9, 64, 5), (33, 2049, 2)])
def test_planes_match_oracle(layout, T, N, F):
    rng = np.random.default_rng(T * 7 + N + F)
    X, M, R, RM = _inputs(rng, T, N, F, np.float32)
    ic, rk = _run("ordered", X, M, R, RM, layout)
    for f in range(F):
        want_ic = c_oracle.rolling_ic(np.ascontiguousarray(X[..., f]), R, M, RM, 30)
        assert np.array_equal(ic[f], want_ic, equal_nan=True), f
        want_rk = c_oracle.rank_rows(X[..., f], M)
        assert np.array_equal(rk[f].view(np.uint32), want_rk.view(np.uint32)), f


def test_planes_f64_ic_match_oracle():
    rng = np.random.default_rng(99)
    T, N, F = 66, 3000, 3
    X, M, R, RM = _inputs(rng, T, N, F, np.float64)
    X[..., 1] += rng.standard_normal((T, N)) * 1e-9  # ties broken below f32 resolution
    ic, _ = _run("ordered", X, M, R, RM, "TNF", ranks=False)
    for f in range(F):
        want = c_oracle.rolling_ic(np.ascontiguousarray(X[..., f]), R, M, RM, 30)
        assert np.array_equal(ic[f], want, equal_nan=True), f


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("layout", ["FTN", "TNF"])
@pytest.mark.parametrize("T,N,F", [(64, 700, 2), (65, 5000, 3), (131, 1500, 4), (300, 333, 1),
                                   (63, 900, 2)])
def test_combined_equals_ordered(dtype, layout, T, N, F):
    rng = np.random.default_rng(T + 31 * N + F)
    X, M, R, RM = _inputs(rng, T, N, F, dtype)
    ic1, rk1 = _run("ordered", X, M, R, RM, layout, min_valid=20)
    ic2, rk2 = _run("combined", X, M, R, RM, layout, min_valid=20)
    assert np.array_equal(ic1.view(np.uint64), ic2.view(np.uint64))
    assert np.array_equal(rk1.view(np.uint32), rk2.view(np.uint32))
    assert not np.any(rk2 == -7.0) and not np.any(ic2 == -7.0)


def test_combined_repeats_in_one_workspace():
    """Back-to-back calls in one workspace, a shape change and a clear in
    between all give the same bits."""
    L = capi.lib()
    ws = capi.Workspace()
    outs = []
    for (T, N, F) in [(128, 2000, 3), (96, 700, 2), (128, 2000, 3)]:
        X, M, R, RM = _inputs(np.random.default_rng(T + N), T, N, F, np.float32)
        dX, dM, dR, dRM = _dev(X), _dev(M.view(np.uint8)), _dev(R), _dev(RM.view(np.uint8))
        ic = torch.empty((F, T), dtype=torch.float64, device="cuda:0")
        for rep in range(2):
            rc = L.fc_rank_ic_dev(vp(dX.data_ptr()), 4, F, T, N, 1, N * F, F, vp(dM.data_ptr()), 0,
                                  vp(dR.data_ptr()), vp(dRM.data_ptr()), 30, vp(ic.data_ptr()),
                                  None, ws.h, None)
            assert rc == 0, rc
            torch.cuda.synchronize()
            outs.append(((T, N, F), ic.cpu().numpy().copy()))
        if T == 96:
            L.fc_workspace_clear(ws.h)
    by = {}
    for shape, v in outs:
        by.setdefault(shape, []).append(v)
    for shape, vs in by.items():
        for v in vs[1:]:
            assert np.array_equal(v.view(np.uint64), vs[0].view(np.uint64)), shape
