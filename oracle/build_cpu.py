"""Builds the C restatement ``oracle/pc_cpu.c`` into ``oracle/_bin/libpc_cpu64.so`` (FP64 oracle)
and ``oracle/_bin/libpc_cpu32.so`` (FP32 CPU baseline) with gcc and OpenMP.  Test
infrastructure only."""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "pc_cpu.c")
BIN = os.path.join(HERE, "_bin")

# No -ffast-math: the kernels compare against -inf.  No -march=native: the GPU hosts may run a
# different CPU than the build machine.
FLAGS = ["-O3", "-fno-math-errno", "-fopenmp", "-shared", "-fPIC", "-std=c11"]


def build(verbose: bool = True):
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        raise RuntimeError("no C compiler found")
    os.makedirs(BIN, exist_ok=True)
    for real, name in (("double", "libpc_cpu64.so"), ("float", "libpc_cpu32.so")):
        out = os.path.join(BIN, name)
        tmp = out + ".tmp"
        cmd = [cc] + FLAGS + [f"-DPCC_REAL={real}", SRC, "-o", tmp, "-lm"]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"oracle build failed:\n{' '.join(cmd)}\n{r.stderr}")
        os.replace(tmp, out)
        if verbose:
            print(f"built {out}", file=sys.stderr)


if __name__ == "__main__":
    build()
