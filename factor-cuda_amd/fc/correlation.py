"""fc.factor_corr and fc.stock_corr.

backend='cpu' runs the numpy CPU backend (numpy output); backend='cuda' runs
the MFMA contraction on the GPU and mirrors the input container/device.
The correlation precision of the GPU path is a module setting
(set_corr_precision): 'contract' (default; FP64 sums with the exact fallback,
|dr| <= 1e-12 against the per-pair oracle), 'fp32', or 'bf16x3' (stated error
<= 1e-5 against FP64). The value-domain precondition (max|x| <= 1e150, min
nonzero |x| >= 1e-150 over valid cells) is checked here for float64 GPU
inputs and raises ValueError; float32 panels are always inside it
(reference: fc/correlation.py:16-93).
"""
import os

import numpy as np

from . import _device as _dev
from . import _util as u
from ._cpu_core import np_factor_corr, np_stock_corr

_PRECISIONS = ("contract", "fp32", "bf16x3")
_precision = os.environ.get("FC_CORR_PRECISION", "contract")
if _precision not in _PRECISIONS:
    _precision = "contract"


def set_corr_precision(precision):
    """Select the GPU correlation tier: 'contract', 'fp32' or 'bf16x3'."""
    global _precision
    if precision not in _PRECISIONS:
        raise ValueError(f"precision must be one of {_PRECISIONS}, got {precision!r}")
    _precision = precision


def get_corr_precision():
    return _precision


_DOMAIN_MSG = "corr 数值域外：max|x|>1e150 或 min 非零|x|<1e-150"


def _gpu_domain_check(arr, mask):
    if arr.dtype == np.float32:
        return
    a = np.asarray(arr, dtype=np.float64)
    if a.ndim == 3:
        T, N, F = a.shape
        flat = a.reshape(T * N, F)
        valid = np.isfinite(flat)
        if mask is not None:
            valid &= np.asarray(mask, dtype=bool).reshape(T * N, 1)
    else:
        flat = a.reshape(-1)
        valid = np.isfinite(flat)
        if mask is not None:
            valid &= np.asarray(mask, dtype=bool).reshape(-1)
    if not u.in_corr_domain(flat, valid):
        raise ValueError(_DOMAIN_MSG)


def factor_corr(data, mask=None, names=None, backend="cpu"):
    if backend not in ("cpu", "cuda"):
        raise ValueError(f"backend must be 'cpu' or 'cuda', got {backend!r}")
    dk, data = _dev.as_tensor(data)
    if backend == "cuda" and _dev.on_device(data) is not None and data.dim() == 3 \
            and data.numel() > 0:
        # GPU tensor: correlated in place, (F,F) result stays on its device
        u.check_tensor(data, name="data", ndim=3, dlpack=dk == "dlpack")
        T, N, F = data.shape
        if names is not None and len(names) != F:
            raise ValueError(f"names length {len(names)} must equal F={F}")
        mask = u.mask_must_match_device(mask, u._device_of(data), "mask")
        if mask is not None:
            _dev.mask_bytes(mask, T, N, "mask")
        if F > 128:
            raise ValueError("F > 128 (factor pair grid cap, HG-2 frozen)")
        if data.dtype == u._torch.float64 and not _dev.domain_ok(data.detach(), mask):
            raise ValueError(_DOMAIN_MSG)
        return _dev.factor_corr(data.detach(), mask, _precision)
    kind, device, arr = u.to_numpy(data, name="data", ndim=3, dtypes="f32f64", kind=dk)
    T, N, F = arr.shape
    if names is not None and len(names) != F:
        raise ValueError(f"names length {len(names)} must equal F={F}")
    if backend == "cuda":
        u.require_cuda("factor_corr backend='cuda'")
        mask = u.mask_must_match_device(mask, device, "mask")
        m = u.check_mask(mask, name="mask", T=T, N=N)
        if F > 128:
            raise ValueError("F > 128 (factor pair grid cap, HG-2 frozen)")
        _gpu_domain_check(arr, m)
        u.sync_entry(device)
        res = u.fcp().factor_corr_f64(arr, m, precision=_precision)
        return u.make_output(res, kind=kind, device=device, mode="corr_cuda")
    m = u.check_mask(mask, name="mask", T=T, N=N)
    return np_factor_corr(arr, m)


def stock_corr(data, mask=None, backend="cpu"):
    if backend not in ("cpu", "cuda"):
        raise ValueError(f"backend must be 'cpu' or 'cuda', got {backend!r}")
    dk, data = _dev.as_tensor(data)
    if backend == "cuda" and _dev.on_device(data) is not None and data.dim() == 2 \
            and data.numel() > 0:
        # GPU tensor: correlated in place, (N,N) result stays on its device
        u.check_tensor(data, name="data", ndim=2, dlpack=dk == "dlpack")
        T, N = data.shape
        mask = u.mask_must_match_device(mask, u._device_of(data), "mask")
        if mask is not None:
            _dev.mask_bytes(mask, T, N, "mask")
        if int(N) * N > (1 << 31) - 1:
            raise ValueError("N*N exceeds INT32_MAX (output grid cap, HG-2 frozen; N <= 46340)")
        if data.dtype == u._torch.float64 and not _dev.domain_ok(data.detach(), mask):
            raise ValueError(_DOMAIN_MSG)
        return _dev.stock_corr(data.detach(), mask, _precision)
    kind, device, arr = u.to_numpy(data, name="data", ndim=2, dtypes="f32f64", kind=dk)
    T, N = arr.shape
    if backend == "cuda":
        u.require_cuda("stock_corr backend='cuda'")
        mask = u.mask_must_match_device(mask, device, "mask")
        m = u.check_mask(mask, name="mask", T=T, N=N)
        if int(N) * N > (1 << 31) - 1:
            raise ValueError("N*N exceeds INT32_MAX (output grid cap, HG-2 frozen; N <= 46340)")
        _gpu_domain_check(arr, m)
        u.sync_entry(device)
        res = u.fcb().stock_corr_f64(arr, m, precision=_precision)
        return u.make_output(res, kind=kind, device=device, mode="corr_cuda")
    m = u.check_mask(mask, name="mask", T=T, N=N)
    return np_stock_corr(arr, m)
