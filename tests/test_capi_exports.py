"""The C-ABI library loads and exports every entry point include/fc_hip.h
declares, and the contract checks that return before any device call give
the reference's status codes (no GPU needed)."""
import ctypes

import numpy as np

import capi


def test_every_declared_symbol_is_exported():
    L = capi.lib()
    declared = capi.header_symbols()
    assert len(declared) >= 18
    for name in declared:
        assert hasattr(L, name), name
    assert capi.lib().fc_version().startswith(b"fc_hip")


def test_contract_status_codes_without_device():
    L = capi.lib()
    x = np.zeros((2, 3), np.float32)
    out = np.zeros((2, 3), np.float32)
    assert L.fc_cs_rank_f32(None, None, 2, 3, 0, capi.ptr(out), None) == -1
    assert L.fc_cs_rank_f32(capi.ptr(x), None, 0, 3, 0, capi.ptr(out), None) == -1
    assert L.fc_cs_rank_f32(capi.ptr(x), None, 1, (1 << 24) + 1, 0, capi.ptr(out), None) == -2
    assert L.fc_cs_rank_f32(capi.ptr(x), None, 1 << 16, 1 << 16, 0, capi.ptr(out), None) == -3
    ic = np.zeros(2)
    assert L.fc_rolling_ic(capi.ptr(x), capi.ptr(x), 4, None, None, 2, 3, 1, capi.ptr(ic),
                           None, None, None) == -4
    o = np.zeros((3, 3))
    assert L.fc_stock_corr(capi.ptr(x), 4, None, 2, 3, 9, capi.ptr(o), None) == -1
    assert L.fc_stock_corr(capi.ptr(x), 4, None, 1, 50000, 0, capi.ptr(o), None) == -3
    assert L.fc_factor_corr(capi.ptr(x), 4, None, 1, 1, 129, 0, capi.ptr(o)) == -3
    st = (ctypes.c_int * 4)()
    outs = (ctypes.c_void_p * 4)(*[out.ctypes.data] * 4)
    # a masked group without a mask is a contract error; statuses start not-executed
    assert L.fc_parameter_scan_f32(capi.ptr(x), None, 2, 3, outs, st, None, None, None, None) == -1
    assert list(st) == [-100] * 4


def test_workspace_lifecycle_without_device():
    ws = capi.Workspace()
    capi.lib().fc_workspace_clear(ws.h)
    capi.lib().fc_workspace_clear(ws.h)   # idempotent
    s = capi.mem_stats()
    assert s.live_bytes >= 0 and s.peak_bytes >= s.live_bytes
