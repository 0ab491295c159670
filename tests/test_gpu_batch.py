"""GPU tests of the batched multi-panel schedule (fc_batch_* in
include/fc_hip.h; switched off, so only the argument contract runs): every panel's masked ascending ranks bitwise against the
oracle, its Spearman IC exactly against the C oracle's integer IC, its
factor_corr against the single-panel device call bitwise and against the
FP64 pairwise oracle within the tier's bound; graph replays interleaved with
eager work match the eager run; tuning leaves every output bit unchanged."""
import ctypes

import numpy as np
import pytest

import capi
from oracle import c_oracle, exact

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _panels(rng, shapes, F, dev):
    host, dev_bufs = [], []
    for (T, N) in shapes:
        C = np.corrcoef(rng.standard_normal((F, 3 * F)))
        L = np.linalg.cholesky(C)
        X = np.round(rng.standard_normal((T, N, F)) @ L.T, 3).astype(np.float32)
        p = rng.uniform(0.6, 1.0, size=T)
        M = rng.random((T, N)) < p[:, None]
        R = (0.05 * X[..., 0] + rng.standard_normal((T, N))).astype(np.float32)
        RM = rng.random((T, N)) < 0.95
        host.append((X, M, R, RM))
        d = dict(X=torch.from_numpy(X).to(dev), M=torch.from_numpy(M.view(np.uint8)).to(dev),
                 R=torch.from_numpy(R).to(dev), RM=torch.from_numpy(RM.view(np.uint8)).to(dev),
                 rank=torch.empty((F, T, N), dtype=torch.float32, device=dev),
                 ic=torch.empty((F, T), dtype=torch.float64, device=dev),
                 corr=torch.empty((F, F), dtype=torch.float64, device=dev))
        dev_bufs.append(d)
    return host, dev_bufs


def _desc(dev_bufs, min_valid):
    arr = (capi.Panel * len(dev_bufs))()
    for k, d in enumerate(dev_bufs):
        T, N, F = d["X"].shape
        arr[k] = capi.Panel(d["X"].data_ptr(), d["M"].data_ptr(), d["R"].data_ptr(),
                            d["RM"].data_ptr(), T, N, F, min_valid, d["rank"].data_ptr(),
                            d["ic"].data_ptr(), d["corr"].data_ptr())
    return arr


def _snapshot(dev_bufs):
    torch.cuda.synchronize()
    return [(d["rank"].cpu().numpy().copy(), d["ic"].cpu().numpy().copy(),
             d["corr"].cpu().numpy().copy()) for d in dev_bufs]


def _same(a, b):
    for (r1, i1, c1), (r2, i2, c2) in zip(a, b):
        assert np.array_equal(r1.view(np.uint32), r2.view(np.uint32))
        assert np.array_equal(i1.view(np.uint64), i2.view(np.uint64))
        assert np.array_equal(c1.view(np.uint64), c2.view(np.uint64))


@pytest.mark.skip(reason="fc_batch_create is switched off (returns 801): a run of this test "
                  "ended the process with a host segmentation fault inside the autotuner, and "
                  "the 64-panel bench run with a glibc heap abort; an AddressSanitizer replay "
                  "of the call sequence on the CPU found no cause, so the batch stays off the GPU")
@pytest.mark.parametrize("precision", [0, 2])
def test_batch_panels_match_oracle_and_single_calls(precision):
    dev = torch.device("cuda:0")
    rng = np.random.default_rng(40 + precision)
    shapes = [(60, 300), (97, 1500), (33, 5000), (140, 777), (61, 2049), (8, 4096), (75, 64)]
    F = 4
    host, bufs = _panels(rng, shapes, F, dev)
    L = capi.lib()
    arr = _desc(bufs, 30)
    b = ctypes.c_void_p()
    assert L.fc_batch_create(arr, len(bufs), precision, ctypes.byref(b)) == 0
    try:
        assert L.fc_batch_run(b, None) == 0
        first = _snapshot(bufs)
        for (X, M, R, RM), (rank, ic, corr) in zip(host, first):
            for f in range(F):
                ref = c_oracle.rank_rows(X[..., f], M)
                assert np.array_equal(rank[f].view(np.uint32), ref.view(np.uint32))
                icr = c_oracle.rolling_ic(np.ascontiguousarray(X[..., f]), R, M, RM, 30)
                assert np.array_equal(ic[f], icr, equal_nan=True)
            Xf = X.reshape(-1, F).astype(np.float64)
            m = np.repeat(M.reshape(-1, 1), F, axis=1)
            ref = exact.corr_matrix_pairwise(Xf, m)
            tol = 1e-12 if precision == 0 else 1e-5
            ok = ~np.isnan(ref)
            assert np.array_equal(np.isnan(corr), np.isnan(ref))
            assert np.abs(corr[ok] - ref[ok]).max() <= tol
        # the batch equals the single-panel device path bitwise
        ws = capi.Workspace()
        for d, (_, _, corr) in zip(bufs, first):
            T, N, F_ = d["X"].shape
            o = torch.empty((F_, F_), dtype=torch.float64, device=dev)
            rc = L.fc_factor_corr_dev(ctypes.c_void_p(d["X"].data_ptr()),
                                      ctypes.c_void_p(d["M"].data_ptr()), T, N, F_, precision,
                                      ctypes.c_void_p(o.data_ptr()), ws.h, None)
            assert rc == 0
            torch.cuda.synchronize()
            assert np.array_equal(o.cpu().numpy().view(np.uint64), corr.view(np.uint64))
        # graph replays with eager work and syncs in between repeat the result
        for _ in range(3):
            for d in bufs:
                d["rank"].fill_(-1.0)
                d["ic"].fill_(-1.0)
                d["corr"].fill_(-1.0)
            torch.cuda.synchronize()
            assert L.fc_batch_run(b, None) == 0
            _same(first, _snapshot(bufs))
        # every tuning candidate leaves the outputs bitwise unchanged
        assert L.fc_batch_tune(b, 1, None) == 801  # the timing autotuner is switched off
        s, br = ctypes.c_int(), ctypes.c_int()
        assert L.fc_batch_get_tuning(b, ctypes.byref(s), ctypes.byref(br)) == 0
        assert s.value in (0, 1, 2) and 1 <= br.value <= 4
        for sched in (0, 1, 2):
            for nb in (1, 2, 4):
                assert L.fc_batch_set_tuning(b, sched, nb) == 0
                assert L.fc_batch_run(b, None) == 0
                _same(first, _snapshot(bufs))
        # one panel outside the graph
        assert L.fc_batch_run_panel(b, 2, None) == 0
        _same(first[2:3], _snapshot(bufs[2:3]))
    finally:
        L.fc_batch_destroy(b)


def test_batch_argument_contract():
    L = capi.lib()
    b = ctypes.c_void_p()
    assert L.fc_batch_create(None, 1, 0, ctypes.byref(b)) == -1
    arr = (capi.Panel * 1)()
    arr[0] = capi.Panel(None, None, None, None, 10, 10, 2, 30, None, None, None)
    assert L.fc_batch_create(arr, 1, 0, ctypes.byref(b)) == -1
    assert L.fc_batch_run(None, None) == -1
    assert L.fc_batch_set_tuning(None, 0, 1) == -1
    assert L.fc_batch_tune(None, 1, None) == -1
    # valid panels are refused (switched off) before anything is allocated
    buf = np.zeros(64, np.float64)
    a = buf.ctypes.data
    arr[0] = capi.Panel(a, None, a, None, 4, 2, 2, 30, None, a, None)
    assert L.fc_batch_create(arr, 1, 0, ctypes.byref(b)) == 801
    assert not b.value
