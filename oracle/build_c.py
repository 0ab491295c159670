"""Build recipe of the C oracle: oracle/c/*.c -> oracle/_build/liboracle_c.so
(test infrastructure; the build folder is git-ignored)."""
from __future__ import annotations

import pathlib
import subprocess

HERE = pathlib.Path(__file__).resolve().parent
OUT = HERE / "_build"


def build() -> pathlib.Path:
    OUT.mkdir(exist_ok=True)
    lib = OUT / "liboracle_c.so"
    tmp = OUT / "liboracle_c.so.tmp"
    subprocess.run(["gcc", "-O2", "-shared", "-fPIC", "-fno-fast-math", "-ffp-contract=off",
                    "-o", str(tmp), str(HERE / "c" / "rank_oracle.c"), "-lm"], check=True)
    tmp.replace(lib)
    return lib


if __name__ == "__main__":
    print(build())
