"""Builds the CPU oracle (test infrastructure) into oracle/_build/liboracle.so.

The oracle is plain C compiled with -ffp-contract=off -fno-fast-math so that
every float expression rounds exactly once per operation, like the HIP kernels
it checks.  The reference's own kernels are CUDA and need nvcc, thrust and
ATen, none of which exist here, so no oracle/_ref build is attempted.
"""
from __future__ import annotations

import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
OUT_DIR = os.path.join(HERE, "_build")
LIB = os.path.join(OUT_DIR, "liboracle.so")


def build(verbose: bool = True) -> str:
    os.makedirs(OUT_DIR, exist_ok=True)
    src = os.path.join(HERE, "sailor_oracle.c")
    tmp = LIB + ".tmp"
    cmd = ["gcc", "-O2", "-std=c11", "-fPIC", "-shared", "-fopenmp",
           "-ffp-contract=off", "-fno-fast-math", "-fexcess-precision=standard",
           "-o", tmp, src, "-lm"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"oracle build failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"[oracle.build] built {LIB}")
    return LIB


if __name__ == "__main__":
    build()
