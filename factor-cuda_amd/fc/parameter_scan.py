"""fc.parameter_scan: the rank variants of one panel over the axes
direction x mask_mode, computed from a single sort per (date) segment in one
launch. Record schema, axis normalisation, mask override and the group
failure whitelist follow the reference adapter (fc/parameter_scan.py:33-132).
"""
import time

import numpy as np

from . import _util as u
from ._workspace import _WorkspaceCache

_SCAN_CACHE = _WorkspaceCache(lambda: u.fcb().CsRankWorkspace())

_AXES = ("direction", "mask_mode")
_DIR_VALUES = ("ascending", "descending")
_MASK_VALUES = ("masked", "unmasked")
# binding group order: asc-masked 0, asc-unmasked 1, desc-masked 2, desc-unmasked 3
_BINDING_INDEX = {"ascending": {"masked": 0, "unmasked": 1},
                  "descending": {"masked": 2, "unmasked": 3}}
_WHITELIST = (9, 701)  # invalid configuration / launch out of resources


def _normalize_axes(axes):
    if not isinstance(axes, (list, tuple)) or not all(
            isinstance(a, (list, tuple)) and len(a) == 2 for a in axes):
        raise ValueError("axes must be a list of (axis_name, values_list) tuples")
    user = {}
    for name, vals in axes:
        if name not in _AXES:
            raise ValueError(f"unknown scan axis {name!r}; allowed: direction, mask_mode")
        if name in user:
            raise ValueError(f"duplicate axis {name!r}")
        if not isinstance(vals, (list, tuple)) or len(vals) == 0:
            raise ValueError(f"axis {name!r} values must be non-empty")
        if len(set(vals)) != len(vals):
            raise ValueError(f"axis {name!r} contains duplicate values")
        allowed = _DIR_VALUES if name == "direction" else _MASK_VALUES
        if not set(vals) <= set(allowed):
            raise ValueError(f"{name} values must be in {allowed}")
        user[name] = list(vals)
    effective = {"direction": user.get("direction", ["ascending"]),
                 "mask_mode": user.get("mask_mode", ["masked"])}
    combos = [(d, m) for d in effective["direction"] for m in effective["mask_mode"]]
    return effective, combos


def parameter_scan(axes, X, mask=None):
    effective, combos = _normalize_axes(axes)
    _, device, x = u.to_numpy(X, name="X", ndim=2, dtypes="f32f64", downcast_to="float32")
    T, N = x.shape
    u.require_cuda("parameter_scan")
    if N > (1 << 24):
        raise ValueError("N > 2^24 (rank precision cap, HG-2 frozen)")
    if any(mm == "masked" for _, mm in combos):
        m = u.check_mask(mask, name="mask", T=T, N=N, allow_none=False)
    else:
        m = np.ones((T, N), dtype=np.bool_)  # all-unmasked scans ignore the mask
    active = [0, 0, 0, 0]
    for d, mm in combos:
        active[_BINDING_INDEX[d][mm]] = 1
    u.sync_entry(device)
    t0 = time.perf_counter()
    with _SCAN_CACHE.use((T, N)) as ws:
        r = u.fcb().parameter_scan_f32(x, m, return_timing=True, active_groups=active,
                                       workspace=ws)
    elapsed_ms = (time.perf_counter() - t0) * 1e3

    groups = []
    for gi, (d, mm) in enumerate(combos):
        g = _BINDING_INDEX[d][mm]
        st = r["group_status"][g]
        rec = {"group_index": gi, "axis_values": {"direction": d, "mask_mode": mm}}
        if st == 0:
            rec.update(result=r["groups"][g], status="ok", error=None, error_stage=None,
                       time_ms=r["time_ms"][g], time_gpu_ms=r["time_gpu_ms"][g])
        elif st in _WHITELIST:
            rec.update(result=None, status="failed", error=f"group-level CUDA error {st}",
                       error_stage="launch", time_ms=0.0, time_gpu_ms=0.0)
        else:
            raise RuntimeError(f"non-whitelist CUDA status {st} for group {g} "
                               "(must be scan-level, not group-failed)")
        groups.append(rec)
    ok = [gr for gr in groups if gr["status"] == "ok"]
    return {
        "spec": effective,
        "groups": groups,
        "summary": {
            "total_groups": len(combos),
            "n_failed": len(groups) - len(ok),
            "total_time_ms": sum(gr["time_ms"] for gr in ok),
            "total_time_gpu_ms": sum(gr["time_gpu_ms"] for gr in ok),
            "elapsed_ms": elapsed_ms,
        },
    }
