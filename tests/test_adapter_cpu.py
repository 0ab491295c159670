"""fc adapter contract on the CPU: public signatures, strict scalars, dtype
and mask gates, error mapping, the CPU backends against the oracle, the
parameter_scan record schema and group-failure whitelist (driven by a fake
binding from the host, no GPU), the workspace cache, and the pybind11
modules' argument gates (which run before any device call). Mirrors the
reference's contract tests (tests/test_adapter_v1.py F01/F02/F05/F18,
tests/test_ws_cache_v1.py)."""
import inspect
import threading

import numpy as np
import pytest

import fc
from fc import _util as u
from fc import _workspace
from oracle import np_arm

HAS_GPU = u._torch is not None and u._torch.cuda.is_available()


def test_public_signatures_locked():
    assert fc.__all__ == ["cross_sectional_rank", "factor_plane", "factor_corr", "stock_corr",
                          "rolling_ic", "parameter_scan", "__version__"]
    sig = {n: str(inspect.signature(getattr(fc, n))) for n in fc.__all__[:-1]}
    assert sig == {
        "cross_sectional_rank": "(values, mask=None, descending=False)",
        "factor_plane": "(factors, f)",
        "factor_corr": "(data, mask=None, names=None, backend='cpu')",
        "stock_corr": "(data, mask=None, backend='cpu')",
        "rolling_ic": "(factor, forward_returns, factor_mask=None, fwd_mask=None, min_valid=30, "
                      "device=None)",
        "parameter_scan": "(axes, X, mask=None)",
    }
    assert fc.__version__ == "1.1.1"
    assert callable(fc.clear_workspaces)


def test_strict_scalars():
    x = np.zeros((2, 3), np.float32)
    for bad in (1, 0, None, "True", np.bool_(True)):
        with pytest.raises(ValueError):
            fc.cross_sectional_rank(x, None, bad)
    for bad in (True, 30.0, "30"):
        with pytest.raises(ValueError):
            fc.rolling_ic(x, x, min_valid=bad, device="cpu")
    with pytest.raises(ValueError):
        fc.rolling_ic(x, x, min_valid=1, device="cpu")
    with pytest.raises(ValueError):
        fc.rolling_ic(x, x, device="gpu")
    with pytest.raises(ValueError):
        fc.factor_plane(np.zeros((2, 3, 4)), True)


def test_dtype_and_container_gates():
    with pytest.raises(ValueError):
        fc.stock_corr(np.zeros((3, 2), np.int32))
    with pytest.raises(TypeError):
        fc.stock_corr([[1.0, 2.0], [3.0, 4.0]])
    with pytest.raises(ValueError):
        fc.stock_corr(np.zeros((3, 2)), mask=np.ones((3, 2), np.uint8))
    with pytest.raises(ValueError):
        fc.stock_corr(np.zeros((3, 2)), mask=np.ones((2, 3), bool))
    with pytest.raises(ValueError):
        fc.factor_corr(np.zeros((3, 2)))               # not 3-D
    with pytest.raises(ValueError):
        fc.factor_corr(np.zeros((2, 3, 2)), names=["a"])
    with pytest.raises(ValueError):
        fc.stock_corr(np.zeros((3, 2)), backend="gpu")


@pytest.mark.skipif(u._torch is None, reason="torch not importable")
def test_dlpack_single_consumption():
    torch = u._torch
    t = torch.arange(6, dtype=torch.float64).reshape(3, 2)
    cap = torch.utils.dlpack.to_dlpack(t)
    out = fc.stock_corr(cap, None, backend="cpu")
    assert out.shape == (2, 2)
    with pytest.raises(ValueError):
        fc.stock_corr(cap, None, backend="cpu")


def test_factor_plane_copy_semantics():
    F3 = np.arange(24, dtype=np.float64).reshape(2, 3, 4)
    p = fc.factor_plane(F3, 2)
    assert p.flags["C_CONTIGUOUS"] and np.array_equal(p, F3[..., 2])
    p[0, 0] = -1
    assert F3[0, 0, 2] != -1
    with pytest.raises(ValueError):
        fc.factor_plane(F3, 4)
    with pytest.raises(ValueError):
        fc.factor_plane(np.zeros((2, 3)), 1)


@pytest.mark.skipif(HAS_GPU, reason="checks the no-GPU error path")
def test_gpu_ops_raise_without_gpu():
    x = np.zeros((2, 3), np.float32)
    with pytest.raises(RuntimeError):
        fc.cross_sectional_rank(x)
    with pytest.raises(RuntimeError):
        fc.parameter_scan([("direction", ["ascending"])], x, np.ones((2, 3), bool))
    with pytest.raises(RuntimeError):
        fc.stock_corr(x, backend="cuda")
    with pytest.raises(RuntimeError):
        fc.rolling_ic(x, x, device="cuda")
    # device=None without a GPU runs on the CPU and returns numpy
    r = fc.rolling_ic(np.random.rand(3, 40), np.random.rand(3, 40), min_valid=5)
    assert isinstance(r, np.ndarray) and r.shape == (3,)


def test_cpu_backends_match_oracle():
    rng = np.random.default_rng(4)
    X = rng.standard_normal((60, 12)) + rng.uniform(-1, 1, 12)
    M = rng.random(X.shape) < 0.9
    X[0, 0] = np.nan
    np.testing.assert_array_equal(fc.stock_corr(X, M), np_arm.np_stock_corr(X, M))
    F3 = rng.standard_normal((10, 15, 4))
    Mf = rng.random((10, 15)) < 0.9
    np.testing.assert_array_equal(fc.factor_corr(F3, Mf), np_arm.np_factor_corr(F3, Mf))
    f = rng.standard_normal((20, 50))
    r = rng.standard_normal((20, 50))
    np.testing.assert_array_equal(fc.rolling_ic(f, r, None, None, 30, device="cpu"),
                                  np_arm.np_rolling_ic(f, r, None, None, 30))


def test_corr_domain_error_message():
    x = np.array([[1e151, 1.0], [2.0, 3.0], [4.0, 5.0]])
    with pytest.raises(ValueError) as ei:
        fc.stock_corr(x)
    assert "数值域外" in str(ei.value) or "1e150" in str(ei.value)


def test_corr_precision_setting():
    assert fc.get_corr_precision() in ("contract", "fp32", "bf16x3")
    old = fc.get_corr_precision()
    fc.set_corr_precision("bf16x3")
    assert fc.get_corr_precision() == "bf16x3"
    with pytest.raises(ValueError):
        fc.set_corr_precision("fp16")
    fc.set_corr_precision(old)


class _FakeScan:
    """Binding double: CPU-computed groups with chosen statuses."""

    def __init__(self, status):
        self.status = status

    def CsRankWorkspace(self):
        raise RuntimeError("no workspace in the fake")

    def parameter_scan_f32(self, x, m, return_timing, active_groups, workspace=None):
        groups = np_arm.np_parameter_scan(x, m)
        st = [s if a else -100 for s, a in zip(self.status, active_groups)]
        return {"groups": [g if s == 0 else None for g, s in zip(groups, st)],
                "group_status": st, "time_ms": [1.0] * 4, "time_gpu_ms": [0.5] * 4,
                "_elapsed_ms_diag": 4.0}


def _with_fake(monkeypatch, status):
    monkeypatch.setattr(u, "fcb", lambda: _FakeScan(status))
    monkeypatch.setattr(u, "require_cuda", lambda ctx="": None)
    import sys
    sys.modules["fc.parameter_scan"]._SCAN_CACHE.clear()


def test_parameter_scan_schema_and_whitelist(monkeypatch):
    x = np.round(np.random.default_rng(0).standard_normal((3, 3)), 1).astype(np.float32)
    m = np.array([[1, 0, 1], [1, 1, 1], [0, 1, 1]], bool)
    axes = [("mask_mode", ["masked", "unmasked"]), ("direction", ["ascending", "descending"])]
    _with_fake(monkeypatch, [0, 0, 0, 0])
    res = fc.parameter_scan(axes, x, m)
    assert res["spec"] == {"direction": ["ascending", "descending"],
                           "mask_mode": ["masked", "unmasked"]}
    assert [g["axis_values"] for g in res["groups"]] == [
        {"direction": "ascending", "mask_mode": "masked"},
        {"direction": "ascending", "mask_mode": "unmasked"},
        {"direction": "descending", "mask_mode": "masked"},
        {"direction": "descending", "mask_mode": "unmasked"}]
    assert res["summary"]["total_groups"] == 4 and res["summary"]["n_failed"] == 0
    _with_fake(monkeypatch, [0, 9, 701, 0])
    res = fc.parameter_scan(axes, x, m)
    assert [g["status"] for g in res["groups"]] == ["ok", "failed", "failed", "ok"]
    assert res["groups"][1]["error_stage"] == "launch" and res["groups"][1]["result"] is None
    assert res["summary"]["n_failed"] == 2 and res["summary"]["total_time_ms"] == 2.0
    _with_fake(monkeypatch, [0, 5, 0, 0])
    with pytest.raises(RuntimeError):
        fc.parameter_scan(axes, x, m)
    # all-unmasked scans ignore the user mask; masked scans require one
    _with_fake(monkeypatch, [0, 0, 0, 0])
    res = fc.parameter_scan([("mask_mode", ["unmasked"])], x, np.zeros((9,), np.int8))
    assert res["summary"]["total_groups"] == 1
    with pytest.raises(ValueError):
        fc.parameter_scan([("mask_mode", ["masked"])], x, None)
    for bad in ([("speed", ["x"])], [("direction", ["up"])], [("direction", [])],
                [("direction", ["ascending"]), ("direction", ["descending"])],
                [("direction", ["ascending", "ascending"])], "direction"):
        with pytest.raises(ValueError):
            fc.parameter_scan(bad, x, m)


def test_workspace_cache_locking_and_clear():
    made = []

    class WS:
        def __init__(self):
            self.cleared = 0
            made.append(self)

        def clear(self):
            self.cleared += 1

    cache = _workspace._WorkspaceCache(WS)
    with cache.use((3, 4)) as a:
        pass
    with cache.use((3, 4)) as b:
        pass
    assert a is b and len(made) == 1
    with cache.use((5, 4)) as c:
        assert c is not a
    # same-key use serialises; clear waits for the in-flight call
    order = []

    def hold():
        with cache.use((3, 4)):
            order.append("in")
            threading.Event().wait(0.05)
            order.append("out")

    th = threading.Thread(target=hold)
    th.start()
    threading.Event().wait(0.01)
    cache.clear()
    order.append("cleared")
    th.join()
    assert order == ["in", "out", "cleared"]
    assert all(w.cleared == 1 for w in made)
    # a failing factory disables the cache (calls still run uncached)
    bad = _workspace._WorkspaceCache(lambda: (_ for _ in ()).throw(RuntimeError("x")))
    with bad.use((1, 1)) as w:
        assert w is None
    assert not bad.enabled
    bad.clear()
    assert bad.enabled


def test_pybind_gates_before_device():
    fcb = u.fcb()
    fcp = u.fcp()
    x64 = np.zeros((3, 4), np.float64)
    with pytest.raises(TypeError):
        fcb.cs_rank_f32(x64)
    with pytest.raises(TypeError):
        fcb.cs_rank_f32([[1.0]])
    with pytest.raises(ValueError):
        fcb.cs_rank_f32(x64.astype(np.float32), np.zeros((3, 4), np.float64))
    with pytest.raises(ValueError):
        fcb.rolling_ic_f64(x64, x64, None, None, 1)
    with pytest.raises(ValueError):
        fcb.rolling_ic_f64(x64, x64, None, None, 30.5)
    with pytest.raises(ValueError):
        fcb.parameter_scan_f32(x64.astype(np.float32), None)
    with pytest.raises(ValueError):
        fcb.parameter_scan_f32(x64.astype(np.float32), None, active_groups=[1, 1, 1])
    with pytest.raises(ValueError):
        fcb.stock_corr_f64(x64, None, False, "fp16")
    with pytest.raises(ValueError):
        fcp.factor_corr_f64(np.zeros((2, 2, 129)))
    for name in ("cs_rank_f32", "rolling_ic_f64", "stock_corr_f64", "parameter_scan_f32",
                 "CsRankWorkspace", "RollingIcWorkspace"):
        assert hasattr(fcb, name)
    assert hasattr(fcp, "factor_corr_f64")
