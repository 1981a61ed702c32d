"""Build the reference's own standalone C++ potential tests (test infrastructure only).

  oracle/_ref/test_thermal_tables  <- cosmolattice_ext/tests/test_thermal_tables.cpp
  oracle/_ref/vprime_audit         <- cosmolattice_ext/tests/vprime_audit.cpp

Both include cosmolattice_ext/models/thermal_tables.hpp, the reference's C++ assembly of
V, V', V'' (tree + thermal J_B/J_F + Coleman-Weinberg) from a binary J table, and print values
on fixed (phi, T) grids.  They are compiled from the sources where they lie in the reference
checkout (g++ -std=c++14 -O2 -I cosmolattice_ext/models, the command in their headers); nothing
is copied.  The tests feed them a table written from this project's Bessel-series J functions
and compare their V' with tip.potential.vprime_full.  Without the checkout this is a no-op.
"""
from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
OUT = HERE / "_ref"
REF = Path(os.environ.get("TIP_REFERENCE", "/root/reference"))
PROGRAMS = ("test_thermal_tables", "vprime_audit")


def build(verbose: bool = False) -> list:
    src_dir = REF / "cosmolattice_ext" / "tests"
    inc = REF / "cosmolattice_ext" / "models"
    if not (inc / "thermal_tables.hpp").exists():
        return []
    OUT.mkdir(exist_ok=True)
    made = []
    for name in PROGRAMS:
        exe = OUT / name
        cmd = [os.environ.get("CXX", "g++"), "-std=c++14", "-O2", "-I", str(inc),
               str(src_dir / f"{name}.cpp"), "-o", str(exe)]
        if verbose:
            print(" ".join(cmd), flush=True)
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            sys.stderr.write(r.stdout + r.stderr)
            raise RuntimeError(f"reference build of {name} failed")
        made.append(exe)
    return made


if __name__ == "__main__":
    for p in build(verbose=True):
        print(p)
