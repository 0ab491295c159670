"""fc_rank_ic_dev_noalloc (include/fc_hip.h): rank + Spearman IC with the
return order in the library's static device array instead of a workspace.
Its outputs equal fc_rank_ic_dev's bitwise (IC and the masked ascending
ranks) on two calls in a row, the second smaller and in the (T,N,F) layout,
and the calls allocate nothing through the library's allocator; sizes
beyond the array are refused with the size-cap status before any launch."""
import ctypes

import numpy as np
import pytest

import capi

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

vp = ctypes.c_void_p


def _case(rng, F, T, N, dtype, layout):
    X = np.round(rng.standard_normal((F, T, N)), 3).astype(dtype)
    X[0, 1, : N // 4] = np.nan
    p = rng.uniform(0.6, 1.0, size=T)
    M = rng.random((T, N)) < p[:, None]
    R = (0.05 * X[0] + rng.standard_normal((T, N))).astype(dtype)
    RM = rng.random((T, N)) < 0.95
    dev = torch.device("cuda:0")
    if layout == "FTN":
        dX = torch.from_numpy(X).to(dev)
        strides = (T * N, N, 1)
    else:
        dX = torch.from_numpy(np.ascontiguousarray(X.transpose(1, 2, 0))).to(dev)
        strides = (1, N * F, F)
    return dict(F=F, T=T, N=N, vb=X.itemsize, dX=dX, strides=strides,
                dM=torch.from_numpy(M.view(np.uint8)).to(dev),
                dR=torch.from_numpy(R).to(dev), dRM=torch.from_numpy(RM.view(np.uint8)).to(dev))


def _outs(c):
    dev = c["dX"].device
    ic = torch.full((c["F"], c["T"]), -7.0, dtype=torch.float64, device=dev)
    rk = torch.full((c["F"], c["T"], c["N"]), -7.0, dtype=torch.float32, device=dev)
    return ic, rk


def _args(c, ic, rk):
    ps, rs, es = c["strides"]
    return (vp(c["dX"].data_ptr()), c["vb"], c["F"], c["T"], c["N"], ps, rs, es,
            vp(c["dM"].data_ptr()), 0, vp(c["dR"].data_ptr()), vp(c["dRM"].data_ptr()), 30,
            vp(ic.data_ptr()), vp(rk.data_ptr()))


def _with_ws(c):
    ic, rk = _outs(c)
    ws = capi.Workspace()
    rc = capi.lib().fc_rank_ic_dev(*_args(c, ic, rk), ws.h, None)
    assert rc == 0, rc
    torch.cuda.synchronize()
    return ic.cpu().numpy(), rk.cpu().numpy()


def _noalloc(c):
    ic, rk = _outs(c)
    before = capi.mem_stats()
    rc = capi.lib().fc_rank_ic_dev_noalloc(*_args(c, ic, rk), None)
    assert rc == 0, rc
    torch.cuda.synchronize()
    after = capi.mem_stats()
    assert after.n_alloc == before.n_alloc and after.live_bytes == before.live_bytes
    return ic.cpu().numpy(), rk.cpu().numpy()


def test_noalloc_equals_workspace_entry_point_two_calls():
    rng = np.random.default_rng(71)
    cases = [_case(rng, 12, 1218, 5000, np.float32, "FTN"),   # the benchmark shape first
             _case(rng, 3, 211, 777, np.float64, "TNF")]      # then a smaller one
    for c in cases:
        print(f"noalloc F={c['F']} T={c['T']} N={c['N']} vb={c['vb']}", flush=True)
        want_ic, want_rk = _with_ws(c)
        got_ic, got_rk = _noalloc(c)
        assert np.array_equal(got_ic.view(np.uint64), want_ic.view(np.uint64))
        assert np.array_equal(got_rk.view(np.uint32), want_rk.view(np.uint32))
        assert np.isfinite(got_ic).any()


def test_noalloc_size_caps():
    L = capi.lib()
    cap = L.fc_rank_ic_noalloc_scratch_bytes()
    assert cap == 32 << 20
    dev = torch.device("cuda:0")
    d = torch.zeros(16, dtype=torch.float64, device=dev)
    p = vp(d.data_ptr())
    # 4*T*N beyond the array: refused before any launch
    T, N = 2048, 4097
    assert 4 * T * N > cap
    assert L.fc_rank_ic_dev_noalloc(p, 4, 1, T, N, T * N, N, 1, None, 0, p, None, 30, p, None,
                                    None) == -3
    # rows beyond the register-resident kernels need sort buffers: refused
    n_big = L.fc_rank_max_n(4) + 1
    assert L.fc_rank_ic_dev_noalloc(p, 4, 1, 1, n_big, n_big, n_big, 1, None, 0, p, None, 30, p,
                                    None, None) == -3
    assert L.fc_rank_ic_dev_noalloc(p, 4, 1, 4, 8, 32, 8, 1, None, 0, p, None, 1, p, None,
                                    None) == -4
    assert L.fc_rank_ic_dev_noalloc(None, 4, 1, 4, 8, 32, 8, 1, None, 0, p, None, 30, p, None,
                                    None) == -1
