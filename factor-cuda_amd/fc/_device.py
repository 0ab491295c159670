"""Device-resident path of the adapter: torch-HIP tensors (and DLPack
capsules of device memory) go straight to libfc_hip.so's device-pointer
entry points, and the results come back as tensors on the same device.

The reference adapter copies such inputs to host numpy, lets the binding copy
them back to the GPU, and copies the result out and in again
(fc/_util.py:90-100, fc/_util.py:247-258). Here nothing leaves HBM: the
kernels read the caller's tensors in place (strided views of a factor plane
included, where the entry point takes strides), write into a fresh device
tensor, and the call returns after the caller's current stream has drained,
which is the reference's return contract (fc/_util.py:231-245). Calls that
use a cached workspace drain the stream while they still hold its lock, so
a second thread on another stream never reuses scratch that kernels of the
first call are still reading.

Every result is bitwise the one the host path gives for the same values:
the same entry points run behind both, dtype conversions are the exact ones
the host path makes (float64 -> float32 round-to-nearest-even for ranks,
float32 -> float64 for mixed-precision IC), and the correlation checks read
the same column statistics.

This module holds no fallback: when libfc_hip.so is missing a call raises.
"""
from __future__ import annotations

import ctypes
import functools
import os
import threading

from . import _util as u
from ._workspace import _WorkspaceCache

_vp = ctypes.c_void_p
_i32 = ctypes.c_int
_i64 = ctypes.c_longlong

_SIGS = {
    "fc_workspace_create": ([_vp], _i32),
    "fc_workspace_clear": ([_vp], None),
    "fc_workspace_destroy": ([_vp], None),
    "fc_rank_scan_dev": ([_vp, _i32, _i32, _i32, _i64, _i64, _i64, _vp, _i64, _vp, _vp, _vp, _vp,
                          _vp], _i32),
    "fc_rank_ic_dev": ([_vp, _i32, _i32, _i32, _i32, _i64, _i64, _i64, _vp, _i64, _vp, _vp, _i32,
                        _vp, _vp, _vp, _vp], _i32),
    "fc_stock_corr_dev_ex": ([_vp, _i32, _vp, _i32, _i32, _i32, _vp, _vp, _vp, _vp], _i32),
    "fc_factor_corr_dev_ex": ([_vp, _i32, _vp, _i32, _i32, _i32, _i32, _vp, _vp, _vp], _i32),
}

_TIERS = {"contract": 0, "fp32": 1, "bf16x3": 2}

# FC_DEVICE_RESIDENT=0 restores the reference's host round trip for GPU
# tensors (same results, more copies)
enabled = os.environ.get("FC_DEVICE_RESIDENT", "1") != "0"

# number of calls that took the device path, per operation (tests read it)
calls = {"cs_rank": 0, "rolling_ic": 0, "factor_corr": 0, "stock_corr": 0}
_calls_lock = threading.Lock()


@functools.lru_cache(maxsize=1)
def lib():
    path = u.BUILD / "libfc_hip.so"
    try:
        L = ctypes.CDLL(str(path))
    except OSError as e:
        raise RuntimeError(f"{path} not loadable; build it with factor-cuda_amd/build_lib.py") from e
    for name, (args, res) in _SIGS.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = res
    return L


class _DevWorkspace:
    """One fc_workspace handle; clear() releases its device buffers."""

    def __init__(self):
        self.h = _vp()
        rc = lib().fc_workspace_create(ctypes.byref(self.h))
        if rc != 0:
            raise RuntimeError(f"fc_workspace_create failed with error code {rc}")

    def clear(self):
        lib().fc_workspace_clear(self.h)

    def __del__(self):
        try:
            lib().fc_workspace_destroy(self.h)
        except Exception:
            pass


_IC_WS = _WorkspaceCache(_DevWorkspace)
_CORR_WS = _WorkspaceCache(_DevWorkspace)


def as_tensor(x):
    """(kind, value): a DLPack capsule becomes the torch tensor it carries
    (kind 'dlpack'); a torch tensor stays as it is (kind 'torch'); anything
    else is returned untouched (kind None) for the host path to judge."""
    t = u._torch
    if t is None:
        return None, x
    if isinstance(x, t.Tensor):
        return "torch", x
    if type(x).__name__ == "PyCapsule":
        try:
            return "dlpack", t.utils.dlpack.from_dlpack(x)
        except (RuntimeError, TypeError, AttributeError, BufferError):
            return None, x
    return None, x


def on_device(*xs):
    """Index of the one GPU that every tensor in xs (None entries ignored)
    lives on, or None when any of them is not a GPU tensor or they span
    devices."""
    t = u._torch
    if not enabled or t is None or not t.cuda.is_available():
        return None
    dev = None
    for x in xs:
        if x is None:
            continue
        if not isinstance(x, t.Tensor) or not x.is_cuda:
            return None
        if dev is None:
            dev = x.device.index
        elif x.device.index != dev:
            return None
    return dev


def _ptr(x):
    return None if x is None else _vp(x.data_ptr())


def _check(rc, fn):
    if rc == 0:
        return
    if rc < 0:
        raise ValueError(f"{fn} contract error code {rc}")
    raise RuntimeError(f"{fn} failed with error code {rc}")


def _count(op):
    with _calls_lock:
        calls[op] += 1


def mask_bytes(m, T, N, name):
    """A (T,N) bool device mask as contiguous bytes (torch.bool is one byte
    per element, 0 or 1)."""
    t = u._torch
    if m is None:
        return None
    if m.dtype != t.bool:
        raise ValueError(f"{name} must be bool, got {m.dtype}")
    if tuple(m.shape) != (T, N):
        raise ValueError(f"{name} shape must be (T,N) matching input, got {tuple(m.shape)}")
    return m.detach().contiguous()


class _OnDevice:
    """Current device and stream of the call: the entry points allocate
    workspace buffers on the current HIP device and run on torch's current
    stream of that device."""

    def __init__(self, dev):
        self.dev = dev

    def __enter__(self):
        t = u._torch
        self._ctx = t.cuda.device(self.dev)
        self._ctx.__enter__()
        self.stream = t.cuda.current_stream(self.dev)
        return self

    def __exit__(self, *exc):
        try:
            if exc[0] is None:
                self.stream.synchronize()
        finally:
            self._ctx.__exit__(*exc)
        return False


def cs_rank(x, mask, descending):
    """Ranks of a (T,N) float32/float64 device tensor; see
    cross_sectional_rank. A float32 view is read in place through its
    strides; float64 is rounded to float32 first, as the host path does."""
    t = u._torch
    T, N = x.shape
    dev = x.device.index
    m = mask_bytes(mask, T, N, "mask")
    with _OnDevice(dev) as c:
        if x.dtype == t.float64:
            x = x.to(t.float32)  # round-to-nearest-even, as numpy astype
        rs, es = x.stride()
        if min(rs, es) < 0:
            x = x.contiguous()
            rs, es = x.stride()
        out = t.empty((T, N), dtype=t.float32, device=x.device)
        g = (2 if descending else 0) if m is not None else (3 if descending else 1)
        outs = [None] * 4
        outs[g] = out
        rc = lib().fc_rank_scan_dev(_ptr(x), 1, T, N, 0, rs, es, _ptr(m), 0, _ptr(outs[0]),
                                    _ptr(outs[1]), _ptr(outs[2]), _ptr(outs[3]),
                                    _vp(c.stream.cuda_stream))
        _check(rc, "cs_rank_gpu")
    _count("cs_rank")
    return out


def rolling_ic(f, r, fmask, rmask, min_valid):
    """Per-date Spearman IC of two (T,N) device tensors; see rolling_ic.
    Equal dtypes are ranked as received; a float32/float64 mix is upcast to
    float64 on both sides, as the host binding does."""
    t = u._torch
    T, N = f.shape
    dev = f.device.index
    fm = mask_bytes(fmask, T, N, "factor_mask")
    rm = mask_bytes(rmask, T, N, "fwd_mask")
    with _OnDevice(dev) as c:
        if f.dtype != r.dtype:
            f = f.to(t.float64)
            r = r.to(t.float64)
        vb = 4 if f.dtype == t.float32 else 8
        rs, es = f.stride()
        if min(rs, es) < 0:
            f = f.contiguous()
            rs, es = f.stride()
        r = r.contiguous()
        ic = t.empty((T,), dtype=t.float64, device=f.device)
        with _IC_WS.use(("dev", dev)) as ws:
            if ws is None:
                raise RuntimeError("device workspace cache unavailable")
            rc = lib().fc_rank_ic_dev(_ptr(f), vb, 1, T, N, 0, rs, es, _ptr(fm), 0, _ptr(r),
                                      _ptr(rm), min_valid, _ptr(ic), None, ws.h,
                                      _vp(c.stream.cuda_stream))
            _check(rc, "rolling_ic_gpu")
            c.stream.synchronize()  # the kernels read ws: drain before its lock goes
    _count("rolling_ic")
    return ic


def domain_ok(x, mask):
    """max|x| <= 1e150 and min nonzero |x| >= 1e-150 over the finite,
    mask-selected cells of a float64 device tensor ((T,N) or (T,N,F) with a
    (T,N) mask); an empty selection is inside. Max and min are exact, so
    this agrees with the host check cell for cell."""
    t = u._torch
    a = x.abs()
    valid = t.isfinite(x)
    if mask is not None:
        mm = mask if x.dim() == 2 else mask.unsqueeze(-1)
        valid = valid & mm
    amax = t.where(valid, a, t.zeros((), dtype=a.dtype, device=a.device)).max()
    nzmin = t.where(valid & (a != 0), a, t.full((), float("inf"), dtype=a.dtype,
                                                 device=a.device)).min()
    hi, lo = t.stack([amax, nzmin]).tolist()
    return hi <= 1e150 and lo >= 1e-150


def factor_corr(x, mask, precision):
    """(F,F) pooled Pearson matrix of a (T,N,F) device tensor; see
    factor_corr."""
    t = u._torch
    T, N, F = x.shape
    dev = x.device.index
    m = mask_bytes(mask, T, N, "mask")
    with _OnDevice(dev) as c:
        x = x.detach().contiguous()
        vb = 4 if x.dtype == t.float32 else 8
        out = t.empty((F, F), dtype=t.float64, device=x.device)
        with _CORR_WS.use(("factor", dev)) as ws:
            if ws is None:
                raise RuntimeError("device workspace cache unavailable")
            rc = lib().fc_factor_corr_dev_ex(_ptr(x), vb, _ptr(m), T, N, F, _TIERS[precision],
                                             _ptr(out), ws.h, _vp(c.stream.cuda_stream))
            _check(rc, "factor_corr_gpu")
            c.stream.synchronize()  # the kernels use ws scratch: drain before its lock goes
    _count("factor_corr")
    return out


def stock_corr(x, mask, precision):
    """(N,N) pairwise-complete Pearson matrix of a (T,N) device tensor; see
    stock_corr. One small read-back of column statistics picks the path,
    exactly as the host driver does."""
    t = u._torch
    T, N = x.shape
    dev = x.device.index
    m = mask_bytes(mask, T, N, "mask")
    with _OnDevice(dev) as c:
        x = x.detach().contiguous()
        vb = 4 if x.dtype == t.float32 else 8
        out = t.empty((N, N), dtype=t.float64, device=x.device)
        with _CORR_WS.use(("stock", dev)) as ws:
            if ws is None:
                raise RuntimeError("device workspace cache unavailable")
            rc = lib().fc_stock_corr_dev_ex(_ptr(x), vb, _ptr(m), T, N, _TIERS[precision],
                                            _ptr(out), None, ws.h, _vp(c.stream.cuda_stream))
            _check(rc, "stock_corr_gpu")
            c.stream.synchronize()  # the kernels use ws scratch: drain before its lock goes
    _count("stock_corr")
    return out
