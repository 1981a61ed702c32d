"""Build the CPU oracle (test infrastructure only).

  oracle/libtip_oracle.so  <- oracle/tip_oracle.c  (gcc, OpenMP, -ffp-contract=off)

The oracle is the checker of the HIP library and the CPU-baseline leg of the benchmark; the
product never links or calls it.
"""
from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
SRC = HERE / "tip_oracle.c"
OUT = HERE / "libtip_oracle.so"


def build(verbose: bool = False) -> Path:
    tmp = OUT.with_suffix(".so.tmp")
    cmd = [
        os.environ.get("CC", "gcc"), "-O3", "-march=x86-64-v3", "-fopenmp", "-fPIC", "-shared",
        "-ffp-contract=off", "-fno-fast-math", "-std=c11", "-D_GNU_SOURCE", "-Wall",
        str(SRC), "-o", str(tmp), "-lm",
    ]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("oracle build failed")
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
