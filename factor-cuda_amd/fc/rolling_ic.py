"""fc.rolling_ic: per-date Spearman IC of factor vs forward returns on the
pairwise-valid cells. device=None runs on the GPU when one is available and
on the CPU otherwise; output is numpy iff every input is numpy and execution
is on the CPU, else a torch tensor on the factor's device (reference
contract, fc/rolling_ic.py:16-57).
"""
from . import _device as _dev
from . import _util as u
from ._cpu_core import np_rolling_ic
from ._workspace import _WorkspaceCache

_ROLL_CACHE = _WorkspaceCache(lambda: u.fcb().RollingIcWorkspace())


def rolling_ic(factor, forward_returns, factor_mask=None, fwd_mask=None,
               min_valid=30, device=None):
    if device not in (None, "cpu", "cuda"):
        raise ValueError(f"device must be None, 'cpu', or 'cuda', got {device!r}")
    if not isinstance(min_valid, int) or isinstance(min_valid, bool):
        raise ValueError("min_valid must be an integer (bool rejected)")
    if min_valid < 2:
        raise ValueError("min_valid must be >= 2")
    fk, factor = _dev.as_tensor(factor)
    rk, forward_returns = _dev.as_tensor(forward_returns)
    if device != "cpu" and _dev.on_device(factor, forward_returns) is not None:
        # both sides on one GPU: scored in place, IC stays on the factor's device
        u.check_tensor(factor, name="factor", ndim=2, dlpack=fk == "dlpack")
        u.check_tensor(forward_returns, name="forward_returns", ndim=2, dlpack=rk == "dlpack")
        if tuple(factor.shape) != tuple(forward_returns.shape):
            raise ValueError("factor and forward_returns shapes must match (T,N), "
                             f"got {tuple(factor.shape)} vs {tuple(forward_returns.shape)}")
        if factor.numel() > 0:
            dvc = u._device_of(factor)
            factor_mask = u.mask_must_match_device(factor_mask, dvc, "factor_mask")
            fwd_mask = u.mask_must_match_device(fwd_mask, dvc, "fwd_mask")
            return _dev.rolling_ic(factor.detach(), forward_returns.detach(), factor_mask,
                                   fwd_mask, min_valid)
    kf, df, f = u.to_numpy(factor, name="factor", ndim=2, dtypes="f32f64", kind=fk)
    kr, dr, r = u.to_numpy(forward_returns, name="forward_returns", ndim=2, dtypes="f32f64",
                           kind=rk)
    T, N = f.shape
    if r.shape != (T, N):
        raise ValueError("factor and forward_returns shapes must match (T,N), "
                         f"got {f.shape} vs {r.shape}")
    factor_mask = u.mask_must_match_device(factor_mask, df, "factor_mask")
    fwd_mask = u.mask_must_match_device(fwd_mask, df, "fwd_mask")
    fm = u.check_mask(factor_mask, name="factor_mask", T=T, N=N)
    rm = u.check_mask(fwd_mask, name="fwd_mask", T=T, N=N)
    gpu = u._torch is not None and u._torch.cuda.is_available()
    exec_gpu = {"cpu": False, None: gpu, "cuda": True}[device]
    if exec_gpu:
        if not gpu:
            raise RuntimeError("device='cuda' but CUDA unavailable")
        src = df if kf == "torch" else (dr if kr == "torch" else None)
        if src:
            u.sync_entry(src)
        with _ROLL_CACHE.use((T, N)) as ws:
            res = u.fcb().rolling_ic_f64(f, r, fm, rm, min_valid, workspace=ws)
        kind = "torch" if (kf == "torch" or kr == "torch") else "numpy"
        return u.make_output(res, kind=kind, device=df, mode="roll")
    res = np_rolling_ic(f, r, fm, rm, min_valid)
    if kf == "numpy" and kr == "numpy":
        return res
    return u._torch.as_tensor(res, device=u._torch_device(df))
