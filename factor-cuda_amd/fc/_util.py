"""Adapter helpers: container detection and normalisation, dtype and mask
gates, device rules, lazy binding load, entry/return synchronisation and the
output-container mapping.

Behaviour follows the reference adapter's contract (fc/_util.py:39-258):
numpy / torch.Tensor / DLPack inputs; dtype whitelist {float32, float64}
(ValueError otherwise); masks strictly bool (uint8 is a ValueError);
non-containers are TypeErrors; a consumed DLPack capsule is a ValueError;
outputs mirror the input container and device.
"""
from __future__ import annotations

import functools
import importlib
import pathlib
import sys

import numpy as np

try:
    import torch as _torch
except Exception:  # torch is optional for the numpy/CPU paths
    _torch = None

from ._cpu_core import in_corr_domain as _in_corr_domain

_F32 = np.float32
_F64 = np.float64

ROOT = pathlib.Path(__file__).resolve().parent.parent
BUILD = ROOT / "build"


def _torch_device(dev: str):
    if _torch is None:
        raise RuntimeError("torch is required for torch/DLPack inputs (not installed)")
    if dev == "cpu":
        return _torch.device("cpu")
    idx = int(dev.split(":")[1]) if ":" in dev else _torch.cuda.current_device()
    return _torch.device(f"cuda:{idx}")


def require_cuda(context: str = ""):
    """GPU unavailable -> RuntimeError (correlations support backend='cpu',
    rolling_ic supports device=None / 'cpu')."""
    if _torch is None or not _torch.cuda.is_available():
        raise RuntimeError(
            f"CUDA unavailable{f': {context}' if context else ''}; "
            "correlation supports backend='cpu', rolling_ic supports "
            "device=None (auto-CPU)")


def _dtype_ok(dt, dtypes):
    if _torch is not None and isinstance(dt, _torch.dtype):
        allowed = {"f32": (_torch.float32,), "f64": (_torch.float64,)}.get(
            dtypes, (_torch.float32, _torch.float64))
        return dt in allowed
    allowed = {"f32": (_F32,), "f64": (_F64,)}.get(dtypes, (_F32, _F64))
    return dt in allowed


def _device_of(t) -> str:
    return "cpu" if not t.is_cuda else f"cuda:{t.device.index}"


def check_tensor(t, *, name, ndim, dtypes="f32f64", dlpack=False):
    """The dtype and rank gates of to_numpy for a torch tensor (one that came
    out of a DLPack capsule when dlpack=True), without moving it."""
    if not _dtype_ok(t.dtype, dtypes):
        raise ValueError(f"{name} dtype {t.dtype} not in whitelist {{float32,float64}}"
                         + (" (capsule consumed)" if dlpack else ""))
    ndim_ok = t.dim() == ndim if isinstance(ndim, int) else t.dim() in ndim
    if not ndim_ok:
        raise ValueError(f"{name} must be {ndim}-D, got ndim={t.dim()}")


def to_numpy(x, *, name, ndim, dtypes="f32f64", downcast_to=None, kind=None):
    """Normalise an input to a C-contiguous numpy array.
    Returns (kind, device, array) with kind in {'numpy', 'torch', 'dlpack'}
    and device 'cpu' or 'cuda:N' (for output mirroring). kind='dlpack' marks
    a torch tensor already taken out of a DLPack capsule."""
    device = "cpu"
    if kind == "dlpack":
        device = _device_of(x)
        t = x.detach()
        if not _dtype_ok(t.dtype, dtypes):
            raise ValueError(f"{name} dtype {t.dtype} not in whitelist {{float32,float64}} "
                             "(capsule consumed)")
        arr = np.ascontiguousarray(t.contiguous().cpu().numpy())
    elif isinstance(x, np.ndarray):
        kind = "numpy"
        if not _dtype_ok(x.dtype, dtypes):
            raise ValueError(f"{name} dtype {x.dtype} not in whitelist {{float32,float64}}")
        arr = np.ascontiguousarray(x)
    elif _torch is not None and isinstance(x, _torch.Tensor):
        kind = "torch"
        device = _device_of(x)
        x = x.detach()
        if not _dtype_ok(x.dtype, dtypes):
            raise ValueError(f"{name} dtype {x.dtype} not in whitelist {{float32,float64}}")
        arr = x.contiguous().cpu().numpy()
    else:
        if _torch is None:
            raise TypeError(f"{name} must be a numpy array, torch.Tensor, or DLPack capsule "
                            "(torch not installed)")
        try:
            t = _torch.utils.dlpack.from_dlpack(x)
        except (RuntimeError, TypeError, AttributeError, BufferError):
            if type(x).__name__ == "PyCapsule":
                raise ValueError(f"{name}: DLPack capsule already consumed") from None
            raise TypeError(f"{name} must be a numpy array, torch.Tensor, or DLPack capsule") from None
        kind = "dlpack"
        device = _device_of(t)
        t = t.detach()
        if not _dtype_ok(t.dtype, dtypes):
            raise ValueError(f"{name} dtype {t.dtype} not in whitelist {{float32,float64}} "
                             "(capsule consumed)")
        arr = np.ascontiguousarray(t.contiguous().cpu().numpy())
    ndim_ok = arr.ndim == ndim if isinstance(ndim, int) else arr.ndim in ndim
    if not ndim_ok:
        raise ValueError(f"{name} must be {ndim}-D, got ndim={arr.ndim}")
    if downcast_to == "float32" and arr.dtype == _F64:
        with np.errstate(over="ignore"):  # beyond float32 range -> +-inf, an invalid cell
            arr = arr.astype(_F32)  # IEEE round-to-nearest-even
    return kind, device, np.ascontiguousarray(arr)


def check_mask(mask, *, name, T, N, allow_none=True):
    """Strict-bool (T, N) mask -> numpy bool array or None."""
    if mask is None:
        if allow_none:
            return None
        raise ValueError(f"{name} is required")
    if isinstance(mask, np.ndarray):
        if mask.dtype != np.bool_:
            raise ValueError(f"{name} must be bool, got {mask.dtype}")
        arr = np.ascontiguousarray(mask)
    elif _torch is not None and isinstance(mask, _torch.Tensor):
        if mask.dtype != _torch.bool:
            raise ValueError(f"{name} must be bool, got {mask.dtype}")
        arr = mask.detach().cpu().numpy()
    elif _torch is not None:
        try:
            t = _torch.utils.dlpack.from_dlpack(mask)
        except (RuntimeError, TypeError, AttributeError, BufferError):
            raise TypeError(f"{name} must be a bool array or DLPack capsule") from None
        if t.dtype != _torch.bool:
            raise ValueError(f"{name} must be bool, got {t.dtype}")
        arr = t.detach().cpu().numpy()
    else:
        raise TypeError(f"{name} must be a bool numpy array or torch.Tensor")
    if arr.shape != (T, N):
        raise ValueError(f"{name} shape must be (T,N) matching input, got {arr.shape}")
    return np.ascontiguousarray(arr)


def mask_must_match_device(mask, device, name):
    """A mask lives on the values' device (no auto-migration). DLPack masks
    are consumed here to learn their device and returned as tensors."""
    if mask is None:
        return None
    if isinstance(mask, np.ndarray):
        md = "cpu"
    elif _torch is not None and isinstance(mask, _torch.Tensor):
        md = _device_of(mask)
    elif _torch is not None:
        try:
            t = _torch.utils.dlpack.from_dlpack(mask)
        except (RuntimeError, TypeError, AttributeError, BufferError):
            return mask  # consumed: check_mask raises
        mask = t
        md = _device_of(t)
    else:
        md = "cpu"
    if md != device:
        raise ValueError(f"{name} must be on the same device as values ({device}), got {md}")
    return mask


@functools.lru_cache(maxsize=None)
def _load_binding(name: str):
    if str(BUILD) not in sys.path:
        sys.path.insert(0, str(BUILD))
    try:
        return importlib.import_module(name)
    except ImportError as e:
        raise RuntimeError(
            f"GPU binding {name} not found in {BUILD}/; build it with "
            "factor-cuda_amd/build_lib.py") from e


def fcb():
    return _load_binding("factor_cuda_pybind")


def fcp():
    return _load_binding("factor_corr_pybind")


def sync_entry(device: str):
    """Drain the producing stream of a GPU input before the op starts."""
    if _torch is not None and device.startswith("cuda") and _torch.cuda.is_available():
        _torch.cuda.current_stream(int(device.split(":")[1])).synchronize()


def sync_return():
    if _torch is not None and _torch.cuda.is_available():
        _torch.cuda.synchronize()


def in_corr_domain(arr, mask=None):
    return _in_corr_domain(np.asarray(arr, dtype=np.float64), mask)


def make_output(np_result, *, kind, device, mode):
    """'mirror' (numpy -> numpy, torch -> torch on the input device),
    'corr_cpu' (always numpy), 'corr_cuda' (mirror), 'roll' (numpy iff numpy
    input and CPU device, else torch on the factor's device)."""
    if mode == "corr_cpu":
        return np_result
    if mode == "roll":
        if kind == "numpy" and device == "cpu":
            return np_result
        return _torch.as_tensor(np_result, device=_torch_device(device))
    if kind == "numpy":
        return np_result
    return _torch.as_tensor(np_result, device=_torch_device(device))
