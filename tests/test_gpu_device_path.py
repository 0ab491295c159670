"""Device-resident boundary: GPU tensors and DLPack capsules of device memory
go to the device-pointer entry points without a host round trip, and every
result is bitwise the one the host path gives for the same values (the
reference adapter's contract, fc/_util.py:90-100 and fc/_util.py:247-258,
with the copies taken out). Each test also checks that the device path ran
(fc._device.calls) and that outputs stay on the input's device."""
import numpy as np
import pytest

import fc
from fc import _device as dv
from fc import _util as u

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def _precision_reset():
    yield
    fc.set_corr_precision("contract")


def _torch():
    return u._torch


def _bits_equal(a, b):
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b)
    assert a.shape == b.shape and a.dtype == b.dtype
    w = np.uint32 if a.dtype == np.float32 else np.uint64
    return np.array_equal(a.view(w), b.view(w))


def _dev_calls(op):
    return dv.calls[op]


def _panel(rng, T, N, dtype=np.float32, nan=0.05, ties=True):
    x = rng.standard_normal((T, N))
    if ties:
        x = np.round(x, 1)
    x[rng.random((T, N)) < nan] = np.nan
    return x.astype(dtype)


def _to(x, dev="cuda:0"):
    t = _torch()
    return None if x is None else t.from_numpy(np.ascontiguousarray(x)).to(dev)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("masked", [False, True])
@pytest.mark.parametrize("descending", [False, True])
def test_cs_rank_device_equals_host(rng, dtype, masked, descending):
    x = _panel(rng, 37, 1500, dtype)
    if dtype == np.float64:
        # values that round differently in float32, overflow to inf, and sit
        # in float32's subnormal range
        x[0, :8] = [1.0 + 2.0 ** -30, 1.0 + 2.0 ** -24, 3e38, 4e38, -1e300, 1e-40, -1e-42, 5e-46]
    m = rng.random(x.shape) < 0.85 if masked else None
    want = fc.cross_sectional_rank(x, m, descending)
    before = _dev_calls("cs_rank")
    got = fc.cross_sectional_rank(_to(x), _to(m), descending)
    assert _dev_calls("cs_rank") == before + 1
    assert got.is_cuda and got.dtype == _torch().float32
    assert _bits_equal(got.cpu().numpy(), want)


def test_cs_rank_strided_views_read_in_place(rng):
    P = _to(_panel(rng, 25, 900, np.float32, nan=0.1).reshape(25, 300, 3))
    for f in range(3):
        view = P[..., f]  # row stride 900, element stride 3
        assert not view.is_contiguous()
        want = fc.cross_sectional_rank(view.cpu().numpy(), None, False)
        got = fc.cross_sectional_rank(view)
        assert _bits_equal(got.cpu().numpy(), want)
    Xt = _to(_panel(rng, 700, 19, np.float32)).t()  # (19, 700), element stride 19
    got = fc.cross_sectional_rank(Xt, None, True)
    assert _bits_equal(got.cpu().numpy(), fc.cross_sectional_rank(Xt.cpu().numpy(), None, True))
    # a factor plane taken with fc.factor_plane stays on its device
    fp = fc.factor_plane(P.cpu(), 1)
    assert _bits_equal(fc.cross_sectional_rank(P[..., 1]).cpu().numpy(),
                       fc.cross_sectional_rank(fp).numpy())
    assert _bits_equal(fc.cross_sectional_rank(P[..., 1]).cpu().numpy(),
                       fc.cross_sectional_rank(P[..., 1].clone()).cpu().numpy())


def test_cs_rank_dlpack_capsule_and_side_stream(rng):
    t = _torch()
    x = _panel(rng, 11, 4000, np.float32)
    m = rng.random(x.shape) < 0.7
    want = fc.cross_sectional_rank(x, m, False)
    before = _dev_calls("cs_rank")
    got = fc.cross_sectional_rank(t.utils.dlpack.to_dlpack(_to(x)),
                                  t.utils.dlpack.to_dlpack(_to(m)))
    assert _dev_calls("cs_rank") == before + 1
    assert isinstance(got, t.Tensor) and got.is_cuda
    assert _bits_equal(got.cpu().numpy(), want)
    s = t.cuda.Stream()
    xd, md = _to(x), _to(m)
    s.wait_stream(t.cuda.current_stream())
    with t.cuda.stream(s):
        got2 = fc.cross_sectional_rank(xd, md)
    t.cuda.synchronize()
    assert _bits_equal(got2.cpu().numpy(), want)


def test_cs_rank_device_errors(rng):
    t = _torch()
    x = _to(_panel(rng, 4, 50))
    with pytest.raises(ValueError):
        fc.cross_sectional_rank(x, np.ones((4, 50), dtype=bool))  # mask on the host
    with pytest.raises(ValueError):
        fc.cross_sectional_rank(x, t.ones((4, 50), dtype=t.uint8, device=x.device))
    with pytest.raises(ValueError):
        fc.cross_sectional_rank(x, t.ones((4, 51), dtype=t.bool, device=x.device))
    with pytest.raises(ValueError):
        fc.cross_sectional_rank(x.to(t.float16))
    with pytest.raises(ValueError):
        fc.cross_sectional_rank(x.reshape(4, 5, 10))


@pytest.mark.parametrize("fdt,rdt", [(np.float32, np.float32), (np.float64, np.float64),
                                     (np.float32, np.float64), (np.float64, np.float32)])
def test_rolling_ic_device_equals_host(rng, fdt, rdt):
    T, N = 40, 2100
    f = _panel(rng, T, N, fdt)
    r = _panel(rng, T, N, rdt, ties=False)
    fm = rng.random((T, N)) < 0.9
    rm = rng.random((T, N)) < 0.95
    fm[3] = False  # an empty date
    for masks in [(None, None), (fm, None), (fm, rm)]:
        want = fc.rolling_ic(f, r, masks[0], masks[1], min_valid=30, device="cuda")
        before = _dev_calls("rolling_ic")
        got = fc.rolling_ic(_to(f), _to(r), _to(masks[0]), _to(masks[1]), min_valid=30)
        assert _dev_calls("rolling_ic") == before + 1
        assert got.is_cuda and got.dtype == _torch().float64 and tuple(got.shape) == (T,)
        assert _bits_equal(got.cpu().numpy(), np.asarray(want))


def test_rolling_ic_strided_factor_plane(rng):
    T, N, F = 30, 800, 4
    P = _to(_panel(rng, T, N * F, np.float64).reshape(T, N, F))
    r = _panel(rng, T, N, np.float64, ties=False)
    rd = _to(r)
    for f in range(F):
        want = fc.rolling_ic(P[..., f].cpu().numpy(), r, min_valid=30, device="cuda")
        got = fc.rolling_ic(P[..., f], rd, min_valid=30)
        assert _bits_equal(got.cpu().numpy(), np.asarray(want))


def test_rolling_ic_device_cpu_request_and_errors(rng):
    T, N = 12, 300
    f = _panel(rng, T, N)
    r = _panel(rng, T, N, ties=False)
    before = _dev_calls("rolling_ic")
    out = fc.rolling_ic(_to(f), _to(r), min_valid=5, device="cpu")
    assert _dev_calls("rolling_ic") == before  # device='cpu' keeps the CPU path
    assert out.is_cuda  # output on the factor's device, as the reference
    with pytest.raises(ValueError):
        fc.rolling_ic(_to(f), _to(r[:, :-1]), min_valid=5)
    with pytest.raises(ValueError):
        fc.rolling_ic(_to(f), _to(r), factor_mask=np.ones((T, N), dtype=bool), min_valid=5)


@pytest.mark.parametrize("precision", ["contract", "fp32", "bf16x3"])
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_factor_corr_device_equals_host(rng, precision, dtype):
    fc.set_corr_precision(precision)
    T, N, F = 23, 310, 9
    x = rng.standard_normal((T, N, F)).astype(dtype)
    x[..., 2] += 1e3  # a column with a large mean
    x[rng.random((T, N, F)) < 0.03] = np.nan
    m = rng.random((T, N)) < 0.9
    for mask in (None, m):
        want = fc.factor_corr(x, mask, backend="cuda")
        before = _dev_calls("factor_corr")
        got = fc.factor_corr(_to(x), _to(mask), backend="cuda")
        assert _dev_calls("factor_corr") == before + 1
        assert got.is_cuda and got.dtype == _torch().float64 and tuple(got.shape) == (F, F)
        assert _bits_equal(got.cpu().numpy(), want)


def test_factor_corr_device_domain_and_caps(rng):
    t = _torch()
    x = rng.standard_normal((5, 40, 3))
    x[2, 7, 1] = 1e200
    with pytest.raises(ValueError):
        fc.factor_corr(_to(x), backend="cuda")
    x[2, 7, 1] = np.nan  # non-finite cells are outside the selection
    assert fc.factor_corr(_to(x), backend="cuda").is_cuda
    m = np.ones((5, 40), dtype=bool)
    x[2, 7, 1] = 1e-200
    m[2, 7] = False  # masked-off cells too
    assert fc.factor_corr(_to(x), _to(m), backend="cuda").is_cuda
    with pytest.raises(ValueError):
        fc.factor_corr(_to(x), backend="cuda")
    with pytest.raises(ValueError):
        fc.factor_corr(t.zeros((2, 3, 129), dtype=t.float32, device="cuda:0"), backend="cuda")
    with pytest.raises(ValueError):
        fc.factor_corr(_to(x), names=["a"], backend="cuda")


@pytest.mark.parametrize("precision", ["contract", "fp32", "bf16x3"])
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_stock_corr_device_equals_host(rng, precision, dtype):
    fc.set_corr_precision(precision)
    T, N = 120, 700
    full = rng.standard_normal((T, N)).astype(dtype)  # fully valid: bf16x3 full-row path
    holed = full.copy()
    holed[rng.random((T, N)) < 0.05] = np.nan
    m = rng.random((T, N)) < 0.9
    for x, mask in ((full, None), (holed, None), (holed, m)):
        want = fc.stock_corr(x, mask, backend="cuda")
        before = _dev_calls("stock_corr")
        got = fc.stock_corr(_to(x), _to(mask), backend="cuda")
        assert _dev_calls("stock_corr") == before + 1
        assert got.is_cuda and got.dtype == _torch().float64 and tuple(got.shape) == (N, N)
        assert _bits_equal(got.cpu().numpy(), want)


def test_device_path_switch_gives_same_results(rng):
    x = _panel(rng, 9, 600)
    xd = _to(x)
    a = fc.cross_sectional_rank(xd)
    old = dv.enabled
    dv.enabled = False
    try:
        before = _dev_calls("cs_rank")
        b = fc.cross_sectional_rank(xd)
        assert _dev_calls("cs_rank") == before
    finally:
        dv.enabled = old
    assert b.is_cuda
    assert _bits_equal(a.cpu().numpy(), b.cpu().numpy())


def test_device_workspaces_survive_shape_changes_and_clear(rng):
    shapes = [(50, 3000), (7, 200), (50, 3000), (31, 5000)]
    for T, N in shapes:
        f = _panel(rng, T, N)
        r = _panel(rng, T, N, ties=False)
        want = fc.rolling_ic(f, r, min_valid=10, device="cuda")
        got = fc.rolling_ic(_to(f), _to(r), min_valid=10)
        assert _bits_equal(got.cpu().numpy(), np.asarray(want))
        fc.clear_workspaces()
