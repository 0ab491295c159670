"""fc adapter end-to-end on the GPU: the frozen parity manifest executed
through fc with the GPU backends, container/device mirroring for torch
inputs, cached vs uncached workspaces bitwise, repeat-call determinism and
the rolling_ic label fixture (reference tests/test_adapter_v1.py F17/F18,
tests/test_ws_cache_v1.py, tests/test_timeline_no_lookahead_v1.py)."""
import json
import pathlib

import numpy as np
import pytest

import fc
from fc import _util as u
from oracle import c_oracle, exact, np_arm

pytestmark = pytest.mark.gpu
GOLD = pathlib.Path(__file__).resolve().parent / "golden"


def test_frozen_manifest_through_fc_gpu():
    z = np.load(GOLD / "parity_anchors_v1.npz", allow_pickle=False)
    m = json.loads((GOLD / "parity_anchors_v1.manifest.json").read_text())
    for c in m["case_schema"]:
        inputs = [z[k] for k in c["inputs"]]
        mask = z[c["masks"]["mask"]] if (c.get("masks") or {}).get("mask") else None
        op, exp = c["operation"], c["expected"]
        if op == "cs_rank":
            x = inputs[0].reshape(1, -1)
            m2 = None if mask is None else mask.reshape(1, -1)
            got = np.asarray(fc.cross_sectional_rank(x, m2, c["params"]["descending"])).reshape(-1)
            e = np.array([np.nan if v == "nan" else v for v in exp], dtype=np.float32)
            assert np.array_equal(np.isnan(got), np.isnan(e)), c["id"]
            assert np.array_equal(got[~np.isnan(e)], e[~np.isnan(e)]), c["id"]
            assert np.all(got[np.isnan(got)].view(np.uint32) == 0x7FC00000)
        elif op == "rolling_ic":
            f = inputs[0].reshape(1, -1)
            r = inputs[1].reshape(1, -1)
            got = float(np.asarray(fc.rolling_ic(f, r, min_valid=c["params"]["min_valid"],
                                                 device="cuda"))[0])
            if exp == "nan":
                assert np.isnan(got), c["id"]
            else:
                assert abs(got - float(exp)) <= 1e-12, c["id"]
        elif op == "factor_corr":
            if len(inputs) == 1:
                with pytest.raises(ValueError):
                    fc.factor_corr(inputs[0], None, backend="cuda")
                continue
            F3 = np.stack([inputs[0], inputs[1]], axis=1).reshape(1, -1, 2)
            if exp == "ValueError":
                with pytest.raises(ValueError):
                    fc.factor_corr(F3, None, backend="cuda")
                continue
            m2 = None if mask is None else mask.reshape(1, -1)
            got = float(np.asarray(fc.factor_corr(F3, m2, backend="cuda"))[0, 1])
            if exp == "nan":
                assert np.isnan(got), c["id"]
            else:
                assert abs(got - float(exp)) <= 1e-12, c["id"]
        elif op == "parameter_scan":
            res = fc.parameter_scan(c["params"]["axes"], inputs[0], mask)
            assert [g["status"] for g in res["groups"]] == ["ok"] * 4
            refs = np_arm.np_parameter_scan(inputs[0], mask)
            for g, ref in zip(res["groups"], refs):
                assert np.array_equal(g["result"].view(np.uint32), ref.view(np.uint32))


def test_torch_device_mirroring():
    torch = u._torch
    dev = torch.device("cuda:0")
    rng = np.random.default_rng(0)
    X = torch.from_numpy(np.round(rng.standard_normal((5, 64)), 1).astype(np.float32)).to(dev)
    M = torch.from_numpy(rng.random((5, 64)) < 0.8).to(dev)
    r = fc.cross_sectional_rank(X, M)
    assert isinstance(r, torch.Tensor) and r.device == X.device and r.dtype == torch.float32
    ref = np_arm.np_cs_rank(X.cpu().numpy(), M.cpu().numpy(), False)
    assert np.array_equal(r.cpu().numpy().view(np.uint32), ref.view(np.uint32))
    with pytest.raises(ValueError):
        fc.cross_sectional_rank(X, M.cpu())          # mask on another device
    ic = fc.rolling_ic(X, X * 2 + 1, M, None, min_valid=5)
    assert isinstance(ic, torch.Tensor) and ic.device == X.device and ic.dtype == torch.float64
    c = fc.stock_corr(X.double(), None, backend="cuda")
    assert isinstance(c, torch.Tensor) and c.device == X.device
    cn = fc.stock_corr(X.cpu().numpy().astype(np.float64), None, backend="cuda")
    assert isinstance(cn, np.ndarray)


def test_workspace_cached_vs_uncached_bitwise():
    rng = np.random.default_rng(1)
    X = np.round(rng.standard_normal((40, 700)), 2).astype(np.float32)
    M = rng.random(X.shape) < 0.8
    R = rng.standard_normal((40, 700))
    import sys
    caches = [sys.modules["fc.cross_sectional_rank"]._RANK_CACHE,
              sys.modules["fc.rolling_ic"]._ROLL_CACHE]
    a1 = fc.cross_sectional_rank(X, M, True)
    i1 = fc.rolling_ic(X, R, M, None, 30, device="cuda")
    a2 = fc.cross_sectional_rank(X, None, True)     # mask on/off alternation
    for c in caches:
        c.set_enabled(False)
    try:
        b1 = fc.cross_sectional_rank(X, M, True)
        j1 = fc.rolling_ic(X, R, M, None, 30, device="cuda")
    finally:
        for c in caches:
            c.set_enabled(True)
    assert np.array_equal(a1.view(np.uint32), b1.view(np.uint32))
    assert np.array_equal(i1, j1, equal_nan=True)
    assert np.array_equal(a2.view(np.uint32), np_arm.np_cs_rank(X, None, True).view(np.uint32))
    fc.clear_workspaces()
    a3 = fc.cross_sectional_rank(X, M, True)
    assert np.array_equal(a1.view(np.uint32), a3.view(np.uint32))


def test_repeat_calls_bitwise_deterministic():
    rng = np.random.default_rng(2)
    X = rng.standard_normal((300, 120))
    M = rng.random(X.shape) < 0.9
    for prec in ("contract", "fp32", "bf16x3"):
        fc.set_corr_precision(prec)
        try:
            a = fc.stock_corr(X, M, backend="cuda")
            b = fc.stock_corr(X, M, backend="cuda")
        finally:
            fc.set_corr_precision("contract")
        assert np.array_equal(a, b, equal_nan=True), prec
    F3 = rng.standard_normal((30, 40, 7))
    assert np.array_equal(fc.factor_corr(F3, None, backend="cuda"),
                          fc.factor_corr(F3, None, backend="cuda"), equal_nan=True)


def test_label_fixture_gpu():
    z = np.load(GOLD / "rolling_ic_labels_v1.npz", allow_pickle=False)
    factor = -np.log(z["price"])
    fwd = z["forward_returns"]
    got = fc.rolling_ic(factor, fwd, min_valid=3, device="cuda")
    ref = c_oracle.rolling_ic(factor, fwd, None, None, 3)
    assert np.array_equal(np.asarray(got), ref, equal_nan=True)


def test_parameter_scan_gpu_timing_fields():
    rng = np.random.default_rng(3)
    X = np.round(rng.standard_normal((20, 300)), 1).astype(np.float32)
    M = rng.random(X.shape) < 0.7
    res = fc.parameter_scan([("direction", ["ascending", "descending"]),
                             ("mask_mode", ["masked", "unmasked"])], X, M)
    assert res["summary"]["n_failed"] == 0
    assert res["summary"]["total_time_gpu_ms"] > 0.0
    assert res["summary"]["elapsed_ms"] >= 0.0
    refs = np_arm.np_parameter_scan(X, M)
    for g, ref in zip(res["groups"], refs):
        assert np.array_equal(g["result"].view(np.uint32), ref.view(np.uint32))


def test_corr_oracle_direct_gpu():
    """Every pair against the per-pair oracle at 1e-12 (contract tier), as
    the reference's F10 oracle-direct tests."""
    rng = np.random.default_rng(4)
    F3 = rng.standard_normal((60, 15, 4))
    mask = rng.random((60, 15)) < 0.85
    F3[0, 0, 1] = np.nan
    got = np.asarray(fc.factor_corr(F3, mask, backend="cuda"))
    T, N, F = F3.shape
    X = F3.reshape(T * N, F)
    ref = exact.corr_matrix_pairwise(X, np.repeat(mask.reshape(-1, 1), F, axis=1))
    assert np.array_equal(np.isnan(got), np.isnan(ref))
    ok = ~np.isnan(ref)
    assert np.max(np.abs(got[ok] - ref[ok])) <= 1e-12
    X2 = rng.standard_normal((80, 33))
    m2 = rng.random(X2.shape) < 0.8
    got2 = np.asarray(fc.stock_corr(X2, m2, backend="cuda"))
    ref2 = exact.corr_matrix_pairwise(X2, m2)
    assert np.array_equal(np.isnan(got2), np.isnan(ref2))
    ok = ~np.isnan(ref2)
    assert np.max(np.abs(got2[ok] - ref2[ok])) <= 1e-12
