"""Build recipe for the gfx950 library and its pybind11 modules.

Outputs (in-tree, next to this file under build/):
  libfc_hip.so                 HIP kernels + C-ABI host drivers (hipcc, gfx950)
  factor_cuda_pybind.so        pybind11 module (cs_rank / rolling_ic / stock_corr /
                               parameter_scan), calls only the C-ABI
  factor_corr_pybind.so        pybind11 module (factor_corr), calls only the C-ABI

Every invocation recompiles. Usage: python build_lib.py [--jobs N]
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import pathlib
import subprocess
import sys
import sysconfig

HERE = pathlib.Path(__file__).resolve().parent
REPO = HERE.parent
CSRC = HERE / "csrc"
BUILD = HERE / "build"
INCLUDE = REPO / "include"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

HIP_SOURCES = ["rank_kernels.hip", "corr_kernels.hip", "capi.hip", "corr_capi.hip", "batch.hip",
               "wide_rows.hip"]

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    # packed FP32 VALU ops stay off in every HIP translation unit
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    f"-I{INCLUDE}", f"-I{CSRC}",
]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    err = "\n".join(line for line in (r.stderr or "").splitlines()
                    if "packed-fp32-ops" not in line)
    if r.returncode != 0:
        raise RuntimeError(f"command failed ({r.returncode}): {' '.join(cmd)}\n{err}")
    return err


def _pybind_flags():
    import pybind11
    return [f"-I{pybind11.get_include()}", f"-I{sysconfig.get_paths()['include']}"]


def build(jobs: int = 8, verbose: bool = False) -> pathlib.Path:
    BUILD.mkdir(parents=True, exist_ok=True)
    objs = []

    def compile_one(src):
        obj = BUILD / (pathlib.Path(src).stem + ".o")
        _run([HIPCC, *HIP_FLAGS, "-c", str(CSRC / src), "-o", str(obj)])
        return obj

    with cf.ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        objs = list(ex.map(compile_one, HIP_SOURCES))
    lib = BUILD / "libfc_hip.so"
    tmp = BUILD / "libfc_hip.so.tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp),
          *[str(o) for o in objs]])
    os.replace(tmp, lib)

    common = ["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-fvisibility=hidden",
              f"-I{INCLUDE}", *_pybind_flags()]
    link = [f"-L{BUILD}", "-lfc_hip", "-Wl,-rpath,$ORIGIN"]
    mods = [("factor_cuda_pybind", "pybind_factor_cuda.cpp"),
            ("factor_corr_pybind", "pybind_factor_corr.cpp")]

    def build_mod(m):
        name, src = m
        out = BUILD / f"{name}.so"
        tmpm = BUILD / f"{name}.so.tmp"
        _run([*common, str(CSRC / src), "-o", str(tmpm), *link])
        os.replace(tmpm, out)
        return out

    with cf.ThreadPoolExecutor(max_workers=2) as ex:
        list(ex.map(build_mod, mods))
    for o in objs:
        try:
            o.unlink()
        except OSError:
            pass
    if verbose:
        print(f"built {lib}")
    return lib


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=8)
    args = ap.parse_args()
    build(args.jobs, verbose=True)
    sys.exit(0)
