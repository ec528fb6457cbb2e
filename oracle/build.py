"""Build recipe of the CPU oracle: oracle/liboracle.so (g++, OpenMP).

-ffp-contract=off keeps every expression in plain IEEE source order, which
is what the device kernels are compiled with as well.  x86-64-v3 keeps the
library loadable on any current x86 host (the host that runs it may differ from
the one that built it).
"""
from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
SRC = HERE / "bz_oracle.cpp"
LIB = HERE / "liboracle.so"

FLAGS = ["-O3", "-march=x86-64-v3", "-ffp-contract=off", "-fno-fast-math", "-fopenmp",
         "-std=c++17", "-shared", "-fPIC", "-Wall"]


def build(verbose: bool = False) -> Path:
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [os.environ.get("CXX", "g++"), *FLAGS, str(SRC), "-o", str(tmp)]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("oracle build failed")
    os.replace(tmp, LIB)
    return LIB


def ensure_built() -> Path:
    if not LIB.exists() or LIB.stat().st_mtime < SRC.stat().st_mtime:
        if not SRC.exists():
            raise RuntimeError("oracle library missing and its source is absent")
        build()
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
