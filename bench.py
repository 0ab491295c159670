"""Benchmark of the MI355X cross-sectional analysis path.

Headline (value): batched rank + IC pipeline on T=1218 dates x N=5000 stocks
x F=12 factors, inputs resident in HBM: per step, one fc_rank_ic_dev call
orders the forward returns once (fwd_info_bucket_kernel) and ranks every
factor plane (masked ascending cs_rank written out) and scores it against
them (Spearman IC) in one launch (rank_ic_bucket_kernel). value = T*N*F
cells / step time. The roofline figure is the rank_ic launch alone
(fc_rank_ic_ordered_dev), timed on its own. Before the W warm-up steps the
process runs untimed steps for 40 ms of GPU time, so that the timed steps
run at the GPU's steady clocks (clock_warmup in the line).

Also reported: the batched parameter_scan (4 rank variants of all 12 planes
in one launch) and stock_corr at N=5000, T=1218 (masked one-factor market
model) in the bf16x3-MFMA and plain-FP32 forms, with max |error| against an
FP64 reference of the same matrix.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
                       [--no-extra] [--profile-only]
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import pathlib
import sys
import time


def _cpu_threads():
    """Threads for the CPU arms: this process's CPU affinity, capped by an
    OMP_NUM_THREADS already in the environment (the host's CPU share)."""
    n = len(os.sched_getaffinity(0))
    try:
        n = min(n, int(os.environ["OMP_NUM_THREADS"]))
    except (KeyError, ValueError):
        pass
    return max(1, n)


# BLAS threads of the numpy CPU arms, fixed before numpy loads its BLAS;
# values already in the environment are kept
CPU_THREADS = _cpu_threads()
for _v in ("OMP_NUM_THREADS", "OPENBLAS_NUM_THREADS", "MKL_NUM_THREADS"):
    os.environ.setdefault(_v, str(CPU_THREADS))

import numpy as np  # noqa: E402

REPO = pathlib.Path(__file__).resolve().parent
LIB = REPO / "factor-cuda_amd" / "build" / "libfc_hip.so"
METRIC = ("stock_corr N=5000 ms bf16x3-MFMA vs FP32 @≤1e-5; "
          "batched rank+IC cells/s")
HBM_PEAK_GBS = 8000.0
COPY_CAL_BYTES = 512 << 20   # bytes read (and written) by the counter-calibration copy
T_, N_, F_ = 1218, 5000, 12


def _lib():
    if not LIB.exists():
        raise RuntimeError(f"{LIB} not built; run __graft_entry__.build() first")
    L = ctypes.CDLL(str(LIB))
    vp, i32, i64 = ctypes.c_void_p, ctypes.c_int, ctypes.c_longlong
    L.fc_workspace_create.argtypes = [vp]
    L.fc_fwd_order_dev.argtypes = [vp, i32, vp, i32, i32, vp, vp]
    L.fc_rank_ic_ordered_dev.argtypes = [vp, i32, i32, i32, i32, i64, i64, i64, vp, i64, i32,
                                         vp, vp, vp, vp]
    L.fc_rank_ic_dev.argtypes = [vp, i32, i32, i32, i32, i64, i64, i64, vp, i64, vp, vp, i32,
                                 vp, vp, vp, vp]
    L.fc_rank_scan_dev.argtypes = [vp, i32, i32, i32, i64, i64, i64, vp, i64, vp, vp, vp, vp, vp]
    L.fc_stock_corr_dev.argtypes = [vp, vp, i32, i32, i32, vp, vp, vp]
    L.fc_factor_corr_dev.argtypes = [vp, vp, i32, i32, i32, i32, vp, vp, vp]
    for f in ("fc_fwd_order_dev", "fc_rank_ic_ordered_dev", "fc_rank_ic_dev", "fc_rank_scan_dev",
              "fc_stock_corr_dev", "fc_factor_corr_dev", "fc_workspace_create"):
        getattr(L, f).restype = i32
    return L


def make_rank_inputs(seed=1):
    """Batched rank/IC workload: per-date valid fraction p_t ~ U(0.6, 1.0),
    mask ~ Bernoulli(p_t); factor values N(0,1) rounded to 3 decimals (~2 %
    exact ties); forward returns 0.05 * factor0 + N(0,1), return mask
    Bernoulli(0.95)."""
    rng = np.random.default_rng(seed)
    p = rng.uniform(0.6, 1.0, size=T_)
    mask = rng.random((T_, N_)) < p[:, None]
    X = np.round(rng.standard_normal((F_, T_, N_)), 3).astype(np.float32)
    fwd = (0.05 * X[0] + rng.standard_normal((T_, N_))).astype(np.float32)
    fmask = rng.random((T_, N_)) < 0.95
    return X, mask, fwd, fmask


def make_ic_cpu_inputs(seed=0, T=250, N=500):
    """rolling_ic CPU-arm workload, T=250 N=500 f32: X ~ N(0,1), returns
    Y = 0.05*X + N(0,1), mask ~ Bernoulli(0.95), 1 % of X set NaN (draw
    order X, noise, mask, NaN positions); scored with min_valid=30."""
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((T, N)).astype(np.float32)
    Y = (0.05 * X + rng.standard_normal((T, N))).astype(np.float32)
    M = rng.random((T, N)) < 0.95
    X[rng.random((T, N)) < 0.01] = np.nan
    return X, Y, M


def make_stock_inputs(N, T=T_, seed=3):
    """One-factor market model with listing gaps: beta ~ U(0.5,1.5), market
    m_t ~ N(0, 0.01^2), r = beta*m + 0.02*t4; stock i listed from s_i ~
    U{0..600}, plus 1 % random holes."""
    rng = np.random.default_rng(seed)
    beta = rng.uniform(0.5, 1.5, size=N)
    mkt = rng.normal(0.0, 0.01, size=T)
    X = (mkt[:, None] * beta[None, :] + 0.02 * rng.standard_t(4, size=(T, N))).astype(np.float32)
    start = rng.integers(0, 601, size=N)
    mask = (np.arange(T)[:, None] >= start[None, :]) & ~(rng.random((T, N)) < 0.01)
    return X, mask


def algorithmic_bytes_rank_ic(T, N, F):
    # read factor planes 4TNF, write cs_rank 4TNF, read factor mask TN, read
    # return-order info 4TN + counts 4T, write IC 8FT
    return 8 * T * N * F + 5 * T * N + 4 * T + 8 * F * T


TRAFFIC_FILE = REPO / "profiles" / "hbm_traffic_rank_ic_scan_stock_factor_corr.md"


def measured_traffic(kernel="rank_ic_bucket_kernel"):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 counter
    summary (FETCH_SIZE and WRITE_SIZE, each scaled by the same counter's
    reading for a copy of known size). None when the summary is absent."""
    try:
        text = TRAFFIC_FILE.read_text()
    except OSError:
        return None
    for line in text.splitlines():
        if line.startswith(kernel + " "):
            kv = dict(t.split("=") for t in line.split()[1:])
            return (float(kv["fetch_kib"]) * 1024 * float(kv["fetch_scale"])
                    + float(kv["write_kib"]) * 1024 * float(kv["write_scale"]))
    return None


def _stats_ms(samples_s, seed=0):
    """Median, interquartile range and a bootstrap 95 % interval of the
    median (2000 resamples), in ms."""
    a = np.asarray(samples_s, dtype=np.float64) * 1e3
    q1, q3 = np.percentile(a, [25, 75])
    rng = np.random.default_rng(seed)
    boots = np.median(rng.choice(a, size=(2000, a.size), replace=True), axis=1)
    lo, hi = np.percentile(boots, [2.5, 97.5])
    return {"median_ms": float(np.median(a)), "iqr_ms": [float(q1), float(q3)],
            "ci95_median_ms": [float(lo), float(hi)], "samples": int(a.size)}


def _cpu_protocol(fn, budget_s):
    """The reference's CPU timing protocol (benchmarks/perf_bench_v1.py:197-237):
    3 warm-ups + 20 timed samples, 1 + 5 for calls over 5 s. The sample
    count is cut (never below 5) so the samples fit budget_s. Returns the
    statistics and the first call's result."""
    t0 = time.perf_counter()
    out = fn()
    first = time.perf_counter() - t0
    warm, n = (1, 5) if first > 5.0 else (3, 20)
    n = int(max(5, min(n, budget_s // max(first, 1e-9))))
    for _ in range(warm - 1):
        fn()
    ts = []
    for _ in range(n):
        t0 = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t0)
    st = _stats_ms(ts)
    st["warmups"] = warm
    return st, out


def _blas_info():
    try:
        from threadpoolctl import threadpool_info
        return [{"api": d.get("internal_api"), "threads": d.get("num_threads")}
                for d in threadpool_info()]
    except Exception as e:  # reported, never hidden
        return [{"error": repr(e)}]


def cpu_baseline(X, mask, fwd, fmask, budget_s=20.0):
    """Oracle numpy arm (restatement of the reference CPU arm): masked
    ascending cs_rank + rolling_ic of one factor plane per call (planes in
    turn), single-threaded numpy row loops, under the reference protocol
    bounded to budget_s; cells/s from the median call."""
    sys.path.insert(0, str(REPO))
    from oracle import np_arm
    k = [0]

    def one_plane():
        f = k[0] % X.shape[0]
        k[0] += 1
        np_arm.np_cs_rank(X[f], mask, False)
        np_arm.np_rolling_ic(X[f], fwd, mask, fmask, 30)

    st, _ = _cpu_protocol(one_plane, budget_s)
    cells = X.shape[1] * X.shape[2]
    return {"value": cells / (st["median_ms"] * 1e-3), "unit": "cells/s", "cores": 1,
            "kind": "port",
            "sample": f"one factor plane of {X.shape[1]}x{X.shape[2]} per call (np_cs_rank masked "
                      f"asc + np_rolling_ic min_valid=30), {st['warmups']} warm-up + "
                      f"{st['samples']} timed calls",
            "per_plane": st}


def cpu_corr_baseline(gpu, budget_s=30.0):
    """The oracle's restatement of the reference numpy correlation arms
    (np_stock_corr, np_factor_corr; fc/_cpu_core.py:194-269) on the same
    panels as the GPU tiers, BLAS on CPU_THREADS threads, reference protocol
    bounded to budget_s per arm. Beside each: every GPU tier's ms, the
    speedup, and its max |error| against this FP64 result (NaN pattern
    compared on every cell)."""
    sys.path.insert(0, str(REPO))
    from oracle import np_arm
    out = {}
    for key, arm in (("stock_corr_N5000", lambda d: np_arm.np_stock_corr(d["X64"], d["M"])),
                     ("stock_corr_N2000", lambda d: np_arm.np_stock_corr(d["X64"], d["M"])),
                     ("factor_corr_F12", lambda d: np_arm.np_factor_corr(d["X"], d["M"]))):
        d = gpu.get(key)
        if d is None:
            continue
        try:
            st, ref = _cpu_protocol(lambda: arm(d), budget_s)
        except Exception as e:  # reported, never hidden
            out[key] = {"error": repr(e)}
            continue
        res = {"cpu_numpy_arm": st, "threads": CPU_THREADS, "kind": "port"}
        for tier, (ms, mat) in d["tiers"].items():
            nan_g, nan_r = np.isnan(mat), np.isnan(ref)
            ok = ~nan_r & ~nan_g
            res[tier] = {"gpu_ms": ms, "speedup_vs_cpu": st["median_ms"] / ms,
                         "max_abs_err_vs_fp64_oracle": float(np.abs(mat[ok] - ref[ok]).max()),
                         "nan_pattern_mismatches": int((nan_g != nan_r).sum())}
        out[key] = res
    return out


def _median_s(fn, warm=2, reps=7):
    for _ in range(warm):
        fn()
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t0)
    return float(np.median(ts))


def ic_cpu_arm_baseline(L):
    """The rolling_ic CPU-arm workload (make_ic_cpu_inputs): the oracle's
    restatement of the reference numpy arm beside the host-pointer driver
    fc_rolling_ic (H2D + kernels + D2H, as the reference binding), each the
    median of 7 calls after 2 warm-ups, and their largest IC difference."""
    sys.path.insert(0, str(REPO))
    from oracle import np_arm
    X, Y, M = make_ic_cpu_inputs()
    T, N = X.shape
    M8 = np.ascontiguousarray(M, dtype=np.uint8)
    ic = np.empty(T, dtype=np.float64)
    vp = ctypes.c_void_p
    L.fc_rolling_ic.argtypes = [vp, vp, ctypes.c_int, vp, vp, ctypes.c_int, ctypes.c_int,
                                ctypes.c_int, vp, vp, vp, vp]
    L.fc_rolling_ic.restype = ctypes.c_int

    def gpu():
        rc = L.fc_rolling_ic(X.ctypes.data, Y.ctypes.data, 4, M8.ctypes.data, None, T, N, 30,
                             ic.ctypes.data, None, None, None)
        assert rc == 0, rc

    arm_s = _median_s(lambda: np_arm.np_rolling_ic(X, Y, M, None, 30))
    gpu_s = _median_s(gpu)
    arm = np_arm.np_rolling_ic(X, Y, M, None, 30)
    return {"T": T, "N": N, "min_valid": 30, "cpu_arm_s_median": arm_s,
            "gpu_binding_s_median": gpu_s, "max_abs_diff_gpu_vs_arm": float(np.max(np.abs(ic - arm))),
            "cores": 1, "kind": "port"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", action="store_true")
    ap.add_argument("--profile-only", action="store_true",
                    help="for rocprofv3: the timed steps, then 3 batched scans and 3 "
                         "stock_corr calls per tier; prints step times only")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU; run with --gpus 1 "
                          "(multi-GPU sharding is a design note, not built)"}))
        sys.exit(2)

    import torch
    dev = torch.device("cuda:0")
    L = _lib()
    X, mask, fwd, fmask = make_rank_inputs()
    dX = torch.from_numpy(X).to(dev)
    dM = torch.from_numpy(mask.view(np.uint8)).to(dev)
    dR = torch.from_numpy(fwd).to(dev)
    dRM = torch.from_numpy(fmask.view(np.uint8)).to(dev)
    ic = torch.empty((F_, T_), dtype=torch.float64, device=dev)
    rk = torch.empty((F_, T_, N_), dtype=torch.float32, device=dev)
    ws = ctypes.c_void_p()
    assert L.fc_workspace_create(ctypes.byref(ws)) == 0
    stream = torch.cuda.current_stream(dev)
    sp = ctypes.c_void_p(stream.cuda_stream)

    vp = ctypes.c_void_p

    def step():
        rc = L.fc_rank_ic_dev(vp(dX.data_ptr()), 4, F_, T_, N_, T_ * N_, N_, 1, vp(dM.data_ptr()),
                              0, vp(dR.data_ptr()), vp(dRM.data_ptr()), 30, vp(ic.data_ptr()),
                              vp(rk.data_ptr()), ws, sp)
        assert rc == 0, rc

    # the GPU's clocks ramp over the first ~10-20 ms of work (the same kernel
    # runs ~8 % slower at the start of a process than after it): untimed steps
    # until 40 ms of GPU time have passed, then the W warm-up steps
    # (not in --profile-only runs, whose counters are per dispatch)
    clock_steps = 0
    c0 = torch.cuda.Event(enable_timing=True)
    c1 = torch.cuda.Event(enable_timing=True)
    c0.record(stream)
    while not args.profile_only:
        for _ in range(10):
            step()
        clock_steps += 10
        c1.record(stream)
        c1.synchronize()
        if c0.elapsed_time(c1) >= 40.0 or clock_steps >= 400:
            break
    c1.record(stream)
    c1.synchronize()
    clock_ms = c0.elapsed_time(c1)
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(args.steps):
        step()
    e1.record(stream)
    torch.cuda.synchronize()
    step_ms = e0.elapsed_time(e1) / args.steps

    # the dominant kernel alone: rank_ic launches against the order of the
    # same returns, each bracketed by HIP events on the stream it runs on
    rc = L.fc_fwd_order_dev(vp(dR.data_ptr()), 4, vp(dRM.data_ptr()), T_, N_, ws, sp)
    assert rc == 0, rc
    kern_ev = []
    for _ in range(args.steps + 1):
        k0 = torch.cuda.Event(enable_timing=True)
        k1 = torch.cuda.Event(enable_timing=True)
        k0.record(stream)
        rc = L.fc_rank_ic_ordered_dev(vp(dX.data_ptr()), 4, F_, T_, N_, T_ * N_, N_, 1,
                                      vp(dM.data_ptr()), 0, 30, vp(ic.data_ptr()),
                                      vp(rk.data_ptr()), ws, sp)
        assert rc == 0, rc
        k1.record(stream)
        kern_ev.append((k0, k1))
    torch.cuda.synchronize()
    k_ms = float(np.mean([a.elapsed_time(b) for a, b in kern_ev[1:]]))
    cells = T_ * N_ * F_
    value = cells / (step_ms * 1e-3)
    abytes = algorithmic_bytes_rank_ic(T_, N_, F_)
    achieved = abytes / (k_ms * 1e-3) / 1e9
    if args.profile_only:
        print(json.dumps({"step_ms": step_ms, "rank_ic_kernel_ms": k_ms}), flush=True)
        TIME_SETTLE[:] = [0.0, 0.0]
        prof = extras(L, torch, dev, dX, dM, ws, sp, stream, reps=3, e2e=False)
        print(json.dumps(prof), flush=True)
        # calibration for the HBM counters: a plain copy of known size
        a = torch.empty(COPY_CAL_BYTES // 4, dtype=torch.float32, device=dev).fill_(1.0)
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes": COPY_CAL_BYTES}), flush=True)
        return

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "cells/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "clock_warmup": {"steps": clock_steps, "gpu_ms": clock_ms},
        "ms_per_step": step_ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded numpy generators of the benchmark configs)",
        "config": {"workload": "batched rank+IC pipeline: fc_rank_ic_dev (fwd_order + "
                               "rank_ic), T=1218 N=5000 F=12, masked asc cs_rank written + "
                               "Spearman IC min_valid=30",
                   "T": T_, "N": N_, "F": F_},
        "roofline": {"bound": "hbm", "kernel": "rank_ic_bucket_kernel<10, true>", "achieved": achieved,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": measured_traffic(), "algorithmic_bytes": abytes, "kernel_ms": k_ms,
                     "traffic_source": "rocprofv3 FETCH_SIZE/WRITE_SIZE, copy-calibrated, "
                                       "profiles/hbm_traffic_rank_ic_scan_stock_factor_corr.md"},
    }

    kept = {}
    if not args.no_extra:
        out["extra"] = extras(L, torch, dev, dX, dM, ws, sp, stream, keep=kept)

    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(X, mask, fwd, fmask)
        cb = out["cpu_baseline"]
        cb["cores"] = 1
        cb["cpu_model"] = _cpu_model()
        cb["host_cores_available"] = len(os.sched_getaffinity(0))
        cb["blas_threads_env"] = {v: os.environ.get(v) for v in
                                  ("OMP_NUM_THREADS", "OPENBLAS_NUM_THREADS", "MKL_NUM_THREADS")}
        cb["blas_threadpools"] = _blas_info()
        try:
            cb["rolling_ic_T250_N500"] = ic_cpu_arm_baseline(L)
        except Exception as e:  # reported, never hidden
            cb["rolling_ic_T250_N500"] = {"error": repr(e)}
        cb["correlations"] = cpu_corr_baseline({k: v for k, v in kept.items() if v})
    print(json.dumps(out))


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


# (settle_ms, min_timed_ms) of _time; the profile-only mode sets both to 0 so
# that a counter run stays a few dispatches per kernel
TIME_SETTLE = [20.0, 5.0]


def _time(fn, stream, torch, reps=10, warm=2):
    """Mean ms per call of fn over back-to-back calls on stream. After the
    warm-up calls, untimed calls run for about TIME_SETTLE[0] ms of GPU time
    (the clocks ramp over ~15 ms after an idle gap), and the timed calls
    cover at least TIME_SETTLE[1] ms, so a short call's figure is neither
    the ramp's nor the first launch's."""
    settle_ms, min_timed_ms = TIME_SETTLE
    for _ in range(warm):
        fn()
    torch.cuda.synchronize()
    a = torch.cuda.Event(enable_timing=True)
    b = torch.cuda.Event(enable_timing=True)
    a.record(stream)
    fn()
    b.record(stream)
    torch.cuda.synchronize()
    one = max(a.elapsed_time(b), 1e-3)
    for _ in range(min(int(settle_ms / one), 2000)):
        fn()
    reps = max(reps, min(int(min_timed_ms / one) + 1, 2000))
    a.record(stream)
    for _ in range(reps):
        fn()
    b.record(stream)
    torch.cuda.synchronize()
    return a.elapsed_time(b) / reps


def extras(L, torch, dev, dX, dM, ws, sp, stream, reps=10, e2e=True,
           keep=None):
    """The benchmark's other configurations; keep (a dict) receives host
    copies of the correlation panels and matrices for the CPU arms."""
    res = {}
    keep = {} if keep is None else keep
    for k in ("stock_corr_N5000", "stock_corr_N2000", "factor_corr_F12"):
        keep[k] = {}
    outs = [torch.empty((F_, T_, N_), dtype=torch.float32, device=dev) for _ in range(4)]

    def scan():
        rc = L.fc_rank_scan_dev(ctypes.c_void_p(dX.data_ptr()), F_, T_, N_, T_ * N_, N_, 1,
                                ctypes.c_void_p(dM.data_ptr()), 0,
                                *[ctypes.c_void_p(o.data_ptr()) for o in outs], sp)
        assert rc == 0, rc

    ms = _time(scan, stream, torch, reps=reps)
    sb = 20 * T_ * N_ * F_ + T_ * N_
    res["parameter_scan_batched"] = {"ms": ms, "cells_per_s": T_ * N_ * F_ / (ms * 1e-3),
                                     "algorithmic_GBps": sb / (ms * 1e-3) / 1e9,
                                     "hbm_frac": sb / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS}
    del outs
    for key, fn in (("stock_corr", lambda: stock_corr_extra(L, torch, dev, ws, sp, stream,
                                                             reps=min(reps, 5),
                                                             keep=keep["stock_corr_N5000"])),
                    ("stock_corr_N2000", lambda: stock_corr_extra(L, torch, dev, ws, sp, stream,
                                                                   reps=min(reps, 5), N=2000,
                                                                   keep=keep["stock_corr_N2000"])),
                    ("factor_corr", lambda: factor_corr_extra(L, torch, dev, ws, sp, stream,
                                                              reps=min(reps, 5),
                                                              keep=keep["factor_corr_F12"])),
                    ("e2e_pipeline", (lambda: e2e_extra(L, torch, dev, stream, reps=3))
                     if e2e else None)):
        if fn is None:
            res[key] = {"skipped": "profile mode: the kernel summaries cover the benchmark shapes only"}
            continue
        try:
            res[key] = fn()
        except Exception as e:  # reported, never hidden
            res[key] = {"error": repr(e)}
    res["multi_panel"] = {"skipped": "fc_batch_create is switched off (returns 801): runs of "
                                     "the 64-panel graph schedule crashed the host process and "
                                     "no cause is known"}
    return res


def lkj_onion(d, eta, rng):
    """Random d x d correlation matrix with density proportional to
    det(C)^(eta-1): the onion method of Lewandowski, Kurowicka and Joe (2009,
    section 3.2). Grows C one row at a time; the new off-diagonal column is
    A w with A = chol(C_k) and w a uniform direction scaled by sqrt(y),
    y ~ Beta(k/2, beta)."""
    beta = eta + (d - 2) / 2.0
    r12 = 2.0 * rng.beta(beta, beta) - 1.0
    C = np.array([[1.0, r12], [r12, 1.0]])
    for k in range(2, d):
        beta -= 0.5
        y = rng.beta(k / 2.0, beta)
        u = rng.standard_normal(k)
        u /= np.linalg.norm(u)
        z = np.linalg.cholesky(C) @ (np.sqrt(y) * u)
        C = np.block([[C, z[:, None]], [z[None, :], np.ones((1, 1))]])
    return C


def make_factor_inputs(torch, dev, F, T=T_, N=N_, seed=2):
    """factor_corr workload: rows x = L z with L = chol(C), C ~ LKJ(eta=2)
    (numpy default_rng(seed)), z ~ N(0, I_F) drawn on the device
    (torch.Generator seeded with `seed`), mask Bernoulli(0.9) on (T,N).
    Returns the (T,N,F) float32 panel and the (T,N) u8 mask, both resident."""
    rng = np.random.default_rng(seed)
    C = lkj_onion(F, 2.0, rng) if F > 1 else np.ones((1, 1))
    Lc = torch.from_numpy(np.linalg.cholesky(C).astype(np.float32)).to(dev)
    g = torch.Generator(device=dev)
    g.manual_seed(seed)
    R = T * N
    X = torch.empty((R, F), dtype=torch.float32, device=dev)
    step = 1 << 20
    for r0 in range(0, R, step):   # chunked so the draw buffer stays small at F=128
        r1 = min(R, r0 + step)
        z = torch.randn((r1 - r0, F), generator=g, device=dev, dtype=torch.float32)
        X[r0:r1] = z @ Lc.T
    M = (torch.rand((T, N), generator=g, device=dev) < 0.9).to(torch.uint8)
    return X.view(T, N, F), M


def factor_corr_extra(L, torch, dev, ws, sp, stream, reps=5, keep=None):
    """factor_corr at T=1218, N=5000, F in {12, 128}: the FP64 contract tier,
    the plain FP32 form and the bf16x3 form (each with an FP64 finalize);
    the fast forms' deviation from the contract tier's matrix here, their
    errors against the FP64 oracle (F=12) in cpu_baseline.correlations.
    keep (a dict) receives host copies of the F=12 panel and matrices."""
    out = {}
    for F in (12, 128):
        X, M = make_factor_inputs(torch, dev, F)
        R = T_ * N_
        # MFMA work issued by the all-finite (Gram) path: one product per upper
        # factor-block pair of width BW (16 for the contract tier and F <= 16,
        # else 32) over every row; bf16x3 issues 3 bf16 products per pair
        bw = 16 if F <= 16 else 32
        nb = -(-F // bw)
        gram_flops = 2 * R * (nb * (nb + 1) // 2) * bw * bw
        nb16 = -(-F // 16)
        res = {"T": T_, "N": N_, "F": F,
               "algorithmic_bytes": 4 * R * F + R,
               "general_form_fp_macs": R * (F * (F + 1) // 2 + 2 * F * F)}
        mats = {}
        for name, prec in (("fp64_contract", 0), ("fp32", 1), ("bf16x3", 2)):
            o = torch.empty((F, F), dtype=torch.float64, device=dev)

            def run():
                rc = L.fc_factor_corr_dev(ctypes.c_void_p(X.data_ptr()),
                                          ctypes.c_void_p(M.data_ptr()), T_, N_, F, prec,
                                          ctypes.c_void_p(o.data_ptr()), ws, sp)
                assert rc == 0, rc

            ms = _time(run, stream, torch, reps=reps, warm=1)
            mats[name] = o
            if prec == 0:
                flops, peak = 2 * R * (nb16 * (nb16 + 1) // 2) * 256, 78.6   # FP64 MFMA
            elif prec == 1:
                flops, peak = gram_flops, 157.3                            # FP32 MFMA
            else:
                flops, peak = 3 * gram_flops, 2500.0                       # BF16 MFMA
            gbps = res["algorithmic_bytes"] / (ms * 1e-3) / 1e9
            tf = flops / (ms * 1e-3) / 1e12
            res[name] = {"ms": ms, "algorithmic_GBps": gbps, "hbm_frac": gbps / HBM_PEAK_GBS,
                         "mfma_issued_TFLOPs": tf, "mfma_peak_TFLOPs": peak,
                         "mfma_frac": tf / peak}
        ref = mats["fp64_contract"]
        for name in ("fp32", "bf16x3"):
            d = torch.nan_to_num((mats[name] - ref).abs(), nan=0.0)
            res[name]["max_abs_dev_vs_contract_tier"] = float(d.max().item())
            res[name]["nan_pattern_mismatches_vs_contract_tier"] = int(
                (torch.isnan(mats[name]) != torch.isnan(ref)).sum().item())
        out[f"F{F}"] = res
        if keep is not None and F == 12:
            keep.update(X=X.cpu().numpy(), M=M.cpu().numpy().astype(bool),
                        tiers={k: (res[k]["ms"], v.cpu().numpy()) for k, v in mats.items()})
        del X, M
        torch.cuda.empty_cache()
    return out


def stock_time_at_peak_ms(T, N, tier):
    """Device time of one stock_corr call at the MFMA peaks, from the
    algorithmic work with P = N(N+1)/2 pairs: bf16x3 issues
    2(3TP + 4TN^2) bf16 flops (3 split products for the upper-triangle
    cross sums, 2 for each of the full sum and sum-of-squares contractions)
    at 2.5 PFLOP/s plus 2TP int8 ops for the counts at 5.0 POP/s; the FP32
    and FP64 forms 2T(P + 2N^2) flops at 157.3 and 78.6 TFLOP/s (their
    kernels also form the counts on the same MFMA, T*P more MACs, not
    counted here)."""
    P = N * (N + 1) // 2
    if tier == 2:
        return (2.0 * (3 * T * P + 4 * T * N * N) / 2.5e15 + 2.0 * T * P / 5.0e15) * 1e3
    flops = 2.0 * T * (P + 2 * N * N)
    return flops / (157.3e12 if tier == 1 else 78.6e12) * 1e3


def stock_corr_extra(L, torch, dev, ws, sp, stream, reps=5, N=5000, keep=None):
    """stock_corr on the masked market-model panel (make_stock_inputs) in the
    three tiers: device ms, time at the MFMA peak and the fraction of it
    reached, and the fast tiers' deviation from the contract tier's matrix
    (the errors against the FP64 oracle are in cpu_baseline.correlations).
    keep (a dict) receives host copies of the panel and matrices."""
    X, mask = make_stock_inputs(N)
    dX = torch.from_numpy(X).to(dev)
    dM = torch.from_numpy(mask.view(np.uint8)).to(dev)
    N = X.shape[1]
    out = {"T": T_, "N": N}
    mats = {}
    for name, prec in (("fp64_contract", 0), ("fp32", 1), ("bf16x3", 2)):
        o = torch.empty((N, N), dtype=torch.float64, device=dev)

        def run():
            rc = L.fc_stock_corr_dev(ctypes.c_void_p(dX.data_ptr()), ctypes.c_void_p(dM.data_ptr()),
                                     T_, N, prec, ctypes.c_void_p(o.data_ptr()), ws, sp)
            assert rc == 0, rc

        ms = _time(run, stream, torch, reps=reps, warm=1)
        mats[name] = o
        tp = stock_time_at_peak_ms(T_, N, prec)
        out[name] = {"ms": ms, "mfma_time_at_peak_ms": tp, "mfma_frac": tp / ms}
    ref = mats["fp64_contract"]
    for name in ("fp32", "bf16x3"):
        d = (mats[name] - ref).abs()
        d = torch.nan_to_num(d, nan=0.0)
        nanmis = int((torch.isnan(mats[name]) != torch.isnan(ref)).sum().item())
        out[name]["max_abs_dev_vs_contract_tier"] = float(d.max().item())
        out[name]["nan_pattern_mismatches_vs_contract_tier"] = nanmis
    out["speedup_bf16x3_vs_fp32"] = out["fp32"]["ms"] / out["bf16x3"]["ms"]
    if keep is not None:
        keep.update(X64=X.astype(np.float64), M=mask,
                    tiers={k: (out[k]["ms"], v.cpu().numpy()) for k, v in mats.items()})
    return out


def e2e_extra(L, torch, dev, stream, reps=3):
    """The reference's end-to-end F=12 pipeline (benchmarks/poc4_e2e_v1.py:80-101:
    per factor parameter_scan (4 groups) -> rolling_ic of every rank plane
    against the forward returns (fmask = mask, rmask = None, min_valid=30) ->
    factor_corr of the (T,N,F) panel -> (4F, T) IC stack), on synthetic data
    of the benchmark shape (T=1218, N=5000, F=12; the reference used its real
    corpus). Two arms:
      binding  the same calls through this build's drop-in pybind modules on
               host arrays, wall clock incl. H2D/D2H (the reference published
               4.432 s for its GPU arm on an RTX 4060 Laptop,
               benchmarks/results/poc4_e2e_v1.md:15);
      device   inputs resident: one 4-group scan launch over all 12 factors,
               one return ordering, one rank+IC launch over all 48 rank
               planes, factor_corr; the IC stack is that launch's output.
    The two arms' IC stacks and correlation matrices must agree bitwise."""
    sys.path.insert(0, str(REPO / "factor-cuda_amd" / "build"))
    import factor_cuda_pybind as fcb
    import factor_corr_pybind as fcp
    X, mask, fwd, _ = make_rank_inputs()
    F = X.shape[0]
    factors = np.ascontiguousarray(X.transpose(1, 2, 0))   # (T, N, F)

    def binding():
        ranks = []
        for f in range(F):
            ranks.extend(fcb.parameter_scan_f32(factors[:, :, f], mask)["groups"])
        ics = [fcb.rolling_ic_f64(rk, fwd, fmask=mask, rmask=None, min_valid=30,
                                  return_ranks=False) for rk in ranks]
        corr = fcp.factor_corr_f64(factors, mask)
        return np.stack(ics), corr

    binding()  # warm-up
    tb = []
    for _ in range(reps):
        t0 = time.perf_counter()
        ic_b, corr_b = binding()
        tb.append(time.perf_counter() - t0)

    vp = ctypes.c_void_p
    dF = torch.from_numpy(factors).to(dev)
    dM = torch.from_numpy(mask.view(np.uint8)).to(dev)
    dR = torch.from_numpy(fwd).to(dev)
    ranks = torch.empty((4, F, T_, N_), dtype=torch.float32, device=dev)
    ic = torch.empty((4 * F, T_), dtype=torch.float64, device=dev)
    corr = torch.empty((F, F), dtype=torch.float64, device=dev)
    ws = ctypes.c_void_p()
    assert L.fc_workspace_create(ctypes.byref(ws)) == 0
    sp = vp(stream.cuda_stream)
    plane = T_ * N_ * F

    def device():
        rc = L.fc_rank_scan_dev(vp(dF.data_ptr()), F, T_, N_, 1, N_ * F, F, vp(dM.data_ptr()), 0,
                                *[vp(ranks[g].data_ptr()) for g in range(4)], sp)
        assert rc == 0, rc
        rc = L.fc_fwd_order_dev(vp(dR.data_ptr()), 4, None, T_, N_, ws, sp)
        assert rc == 0, rc
        rc = L.fc_rank_ic_ordered_dev(vp(ranks.data_ptr()), 4, 4 * F, T_, N_, T_ * N_, N_, 1,
                                      vp(dM.data_ptr()), 0, 30, vp(ic.data_ptr()), None, ws, sp)
        assert rc == 0, rc
        rc = L.fc_factor_corr_dev(vp(dF.data_ptr()), vp(dM.data_ptr()), T_, N_, F, 0,
                                  vp(corr.data_ptr()), ws, sp)
        assert rc == 0, rc

    ms = _time(device, stream, torch, reps=reps * 3, warm=1)
    # device planes are group-major (g*F + f); the binding stacks per factor (f*4 + g)
    ic_d = np.ascontiguousarray(ic.cpu().numpy().reshape(4, F, T_).transpose(1, 0, 2)
                                .reshape(4 * F, T_))
    corr_d = corr.cpu().numpy()
    return {"T": T_, "N": N_, "F": F, "data": "synthetic (benchmark rank generator); the "
            "reference timed its real corpus of the same shape",
            "binding_s_median": float(np.median(tb)), "binding_s_all": tb,
            "device_ms": ms,
            "reference_gpu_arm_s": 4.432, "reference_numpy_arm_s": 48.89,
            "speedup_binding_vs_reference_gpu": 4.432 / float(np.median(tb)),
            "ic_stack_bitwise_equal": bool(np.array_equal(ic_b.view(np.uint64),
                                                          ic_d.view(np.uint64))),
            "factor_corr_bitwise_equal": bool(np.array_equal(corr_b.view(np.uint64),
                                                             corr_d.view(np.uint64)))}


if __name__ == "__main__":
    main()
