"""fc.cross_sectional_rank and fc.factor_plane.

Signatures and behaviour of the reference adapter
(fc/cross_sectional_rank.py:19-54): stable ordinal ranks 1..K per row on the
float32 tensor (float64 input downcast round-to-nearest-even), invalid cells
quiet NaN 0x7fc00000, GPU only (RuntimeError without a GPU), output mirrors
the input container.
"""
import numpy as np

from . import _device as _dev
from . import _util as u
from ._workspace import _WorkspaceCache

_RANK_CACHE = _WorkspaceCache(lambda: u.fcb().CsRankWorkspace())


def cross_sectional_rank(values, mask=None, descending=False):
    if type(descending) is not bool:
        raise ValueError("descending must be a bool (type(x) is bool), got "
                         f"{type(descending).__name__}")
    vk, values = _dev.as_tensor(values)
    if _dev.on_device(values) is not None:
        # GPU tensor: ranked in place, result stays on its device
        u.check_tensor(values, name="values", ndim=2, dtypes="f32f64", dlpack=vk == "dlpack")
        if values.numel() > 0:
            mask = u.mask_must_match_device(mask, u._device_of(values), "mask")
            return _dev.cs_rank(values.detach(), mask, descending)
    kind, device, x = u.to_numpy(values, name="values", ndim=2, dtypes="f32f64",
                                 downcast_to="float32", kind=vk)
    T, N = x.shape
    u.require_cuda("cross_sectional_rank")
    mask = u.mask_must_match_device(mask, device, "mask")
    m = u.check_mask(mask, name="mask", T=T, N=N)
    u.sync_entry(device)
    with _RANK_CACHE.use((T, N)) as ws:
        res = u.fcb().cs_rank_f32(x, m, descending, workspace=ws)
    return u.make_output(res, kind=kind, device=device, mode="mirror")


def factor_plane(factors, f):
    """Factor f of a (T,N,F) panel (or f == 0 of a (T,N) panel) as a fresh
    C-contiguous (T,N) copy; the container mirrors the input."""
    if not isinstance(f, int) or isinstance(f, bool):
        raise ValueError(f"f must be an integer, got {type(f).__name__}")
    kind, device, arr = u.to_numpy(factors, name="factors", ndim=(2, 3), dtypes="f32f64")
    if arr.ndim == 2:
        if f != 0:
            raise ValueError("ndim=2 input requires f == 0")
        plane = np.array(arr, copy=True, order="C")
    else:
        if not 0 <= f < arr.shape[2]:
            raise ValueError(f"f={f} out of range [0,F={arr.shape[2]})")
        plane = np.ascontiguousarray(arr[..., f])
    return u.make_output(plane, kind=kind, device=device, mode="mirror")
