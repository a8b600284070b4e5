"""Build the CPU oracle (oracle/liboracle.so) from oracle/vr_oracle.c.

TEST INFRASTRUCTURE ONLY: the oracle is the checker used by tests/, by
__graft_entry__.smoke() and by bench.py's cpu_baseline leg.

The reference's own implementation of this path (cl/*.cl, OpenCL C with
image3d_t reads and work-item built-ins) needs an OpenCL runtime that this
image does not have (and no CPU OpenCL device), so no oracle/_ref build
exists; the oracle is pinned by known answers and closed forms instead.
"""
from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
OUT = HERE / "liboracle.so"
FLAGS = ["-O3", "-std=c11", "-fPIC", "-shared", "-ffp-contract=off", "-fno-fast-math", "-fopenmp",
         "-Wall", "-Wextra"]


def build() -> Path:
    tmp = HERE / "liboracle.so.tmp"
    cmd = ["gcc", *FLAGS, str(HERE / "vr_oracle.c"), str(HERE / "vr_mlp.c"), "-o", str(tmp), "-lm"]
    res = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"oracle build failed:\n{res.stdout}")
    if res.stdout.strip():
        print(res.stdout, file=sys.stderr)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build())
