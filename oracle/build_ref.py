"""Build the reference's own C++ into oracle/_ref (test infrastructure only: the checker's
checker).  Nothing is copied into the repository; every program is compiled from the reference
checkout where it lies, and the outputs stay in oracle/_ref/ (git-ignored).

  oracle/_ref/test_thermal_tables  <- cosmolattice_ext/tests/test_thermal_tables.cpp
  oracle/_ref/vprime_audit         <- cosmolattice_ext/tests/vprime_audit.cpp
      Both include cosmolattice_ext/models/thermal_tables.hpp, the reference's C++ assembly of
      V, V', V'' (tree + thermal J_B/J_F + Coleman-Weinberg) from a binary J table, and print
      values on fixed (phi, T) grids (g++ -std=c++14 -O2 -I cosmolattice_ext/models, the
      command in their headers).  The tests feed them a table written from this project's
      Bessel-series J functions and compare their V' with tip.potential.vprime_full.
  oracle/_ref/ref_zn_probe         <- oracle/ref_zn_probe.cpp + thermal_tables.hpp
      The reference's complex-field force vPrimeComponents (thermal_tables.hpp:161-178), the
      C++ statement of the Z_N term of latticeSimComplex_numba.py:2147-2152.
  oracle/_ref/ref_hash_probe       <- oracle/ref_hash_probe.cpp + stochasticrk.h:205-225
      hashMix64 / hashRngPair of cosmolattice_ext/evolvers/stochasticrk.h.  That header needs
      TempLat (not in the checkout), so the two static member functions are taken, unchanged,
      from the checkout into oracle/_ref/stochasticrk_hash.inc at build time and compiled as
      members of a class template (their own declaration context), with no stand-in for any
      TempLat header.

Without the checkout this is a no-op.
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


def _cxx(cmd, verbose):
    if verbose:
        print(" ".join(map(str, cmd)), flush=True)
    r = subprocess.run(list(map(str, cmd)), capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"reference build failed: {' '.join(map(str, cmd))}")


def hash_fragment(header: Path) -> str:
    """The text of hashMix64 and hashRngPair, unchanged, from the reference header: from the
    line declaring hashMix64 through the closing brace of hashRngPair (brace matched)."""
    lines = header.read_text().splitlines(keepends=True)
    start = next(i for i, l in enumerate(lines) if "static uint64_t hashMix64(uint64_t x)" in l)
    out, depth, bodies = [], 0, 0
    for l in lines[start:]:
        out.append(l)
        for ch in l:
            if ch == "{":
                depth += 1
            elif ch == "}":
                depth -= 1
                if depth == 0:
                    bodies += 1
        if bodies == 2:
            break
    text = "".join(out)
    if bodies != 2 or "hashRngPair" not in text:
        raise RuntimeError("hashMix64 / hashRngPair not found in the reference header")
    return text


def build(verbose: bool = False) -> list:
    src_dir = REF / "cosmolattice_ext" / "tests"
    inc = REF / "cosmolattice_ext" / "models"
    if not (inc / "thermal_tables.hpp").exists():
        return []
    OUT.mkdir(exist_ok=True)
    cxx = os.environ.get("CXX", "g++")
    made = []
    for name in PROGRAMS:
        exe = OUT / name
        _cxx([cxx, "-std=c++14", "-O2", "-I", inc, src_dir / f"{name}.cpp", "-o", exe], verbose)
        made.append(exe)
    exe = OUT / "ref_zn_probe"
    _cxx([cxx, "-std=c++14", "-O2", "-ffp-contract=off", "-I", inc, HERE / "ref_zn_probe.cpp",
          "-o", exe], verbose)
    made.append(exe)
    hdr = REF / "cosmolattice_ext" / "evolvers" / "stochasticrk.h"
    if hdr.exists():
        (OUT / "stochasticrk_hash.inc").write_text(hash_fragment(hdr))
        exe = OUT / "ref_hash_probe"
        _cxx([cxx, "-std=c++14", "-O2", "-I", OUT, HERE / "ref_hash_probe.cpp", "-o", exe],
             verbose)
        made.append(exe)
    return made


if __name__ == "__main__":
    for p in build(verbose=True):
        print(p)
