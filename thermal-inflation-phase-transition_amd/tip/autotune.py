"""Launch-geometry autotuner for the step kernels.

The knobs are the ones the kernels expose: for a lattice the kernel (LDS row tiles with 1, 2
or 4 y rows per thread, or flat tiles of 1, 2, 4 or 8 rows of 64 z-sites per block), the
number of x planes a block marches through and the XCD-aware block remap; for an ensemble the
flat-tile rows, planes per work item, remap and the number of persistent blocks.  Each candidate is timed with HIP events on the object's own
stream over a few steps; the fastest is applied and remembered in a JSON cache keyed by the
workload signature, so later runs of the same workload skip the search.

Timing steps advance the object's state; tune before uploading the initial condition or
re-upload it afterwards.
"""
from __future__ import annotations

import itertools
import json
import os
from pathlib import Path
from typing import Dict, Iterable, List, Optional, Tuple

CACHE_PATH = Path(os.environ.get("TIP_AUTOTUNE_CACHE",
                                 Path.home() / ".cache" / "tip" / "autotune.json"))


def _load_cache() -> Dict[str, dict]:
    try:
        return json.loads(CACHE_PATH.read_text())
    except (OSError, ValueError):
        return {}


def _store_cache(key: str, value: dict) -> None:
    try:
        cache = _load_cache()
        cache[key] = value
        CACHE_PATH.parent.mkdir(parents=True, exist_ok=True)
        CACHE_PATH.write_text(json.dumps(cache, indent=1, sort_keys=True))
    except OSError:
        pass  # a read-only home only costs the search next time


def lattice_candidates(complex_field: bool = False) -> List[Tuple[int, int, int, int]]:
    """(rows, planes, xcd, kernel): the LDS row-tile kernel (real field) with 1, 2 or 4 y rows
    per thread, and the flat-tile kernel with 1, 2, 4 or 8 rows of 64 sites per block."""
    out = []
    if not complex_field:
        for ty, cx, xcd in itertools.product((1, 2, 4), (8, 16, 32), (1, 0)):
            out.append((ty, cx, xcd, 1))
    for by, cx, xcd in itertools.product((1, 2, 4, 8), (4, 8, 16, 32), (1, 0)):
        out.append((by, cx, xcd, 2))
    return out


def ensemble_candidates() -> List[Tuple[int, int, int, int, int]]:
    """(rows, planes, xcd, persistent, kernel) over row tiles (kernel 1) and flat tiles (2)."""
    out = []
    for ty, cx, pers in itertools.product((1, 2, 4), (8, 16, 32, 64), (0, 2048)):
        out.append((ty, cx, 1, pers, 1))
    for by, cx, pers in itertools.product((2, 4, 8), (8, 16, 32), (0, 2048)):
        out.append((by, cx, 1, pers, 2))
    return out


def tune_lattice(lat, step_kw: dict, candidates: Optional[Iterable] = None, steps: int = 20,
                 use_cache: bool = True) -> dict:
    """Time each (rows, planes, xcd) geometry on a Lattice and apply the fastest."""
    key = f"lattice2:{lat.shape}:{lat.dtype.name}:{'c' if lat.complex else 'r'}"
    if use_cache:
        hit = _load_cache().get(key)
        if hit:
            lat.set_geometry(*hit["best"])
            return hit
    results = []
    for cand in (candidates or lattice_candidates(lat.complex)):
        lat.set_geometry(*cand)
        lat.step(2, timed=True, **step_kw)
        ms = lat.step(steps, timed=True, **step_kw)
        results.append((ms / steps, tuple(cand)))
    results.sort()
    best = results[0][1]
    lat.set_geometry(*best)
    rec = {"best": list(best), "us_per_step": results[0][0] * 1e3,
           "table": [[list(c), t * 1e3] for t, c in results]}
    _store_cache(key, rec)
    return rec


def tune_ensemble(ens, candidates: Optional[Iterable] = None, steps: int = 10,
                  mode: str = "graph", use_cache: bool = True) -> dict:
    """Time each (rows, planes, xcd, persistent) geometry on an Ensemble and apply the best."""
    key = f"ensemble2:{sorted(set(ens.shapes))}:{len(ens.shapes)}:{mode}"
    if use_cache:
        hit = _load_cache().get(key)
        if hit:
            ens.set_geometry(*hit["best"])
            return hit
    results = []
    for cand in (candidates or ensemble_candidates()):
        ens.set_geometry(*cand)
        ens.step(steps, mode=mode, timed=True)  # capture / warm
        ms = ens.step(steps, mode=mode, timed=True)
        results.append((ms / steps, tuple(cand)))
    results.sort()
    best = results[0][1]
    ens.set_geometry(*best)
    rec = {"best": list(best), "us_per_step": results[0][0] * 1e3,
           "table": [[list(c), t * 1e3] for t, c in results]}
    _store_cache(key, rec)
    return rec


def tune_ensemble_schedule(ens, candidates: Iterable[int] = (0, 96 << 20, 160 << 20, 224 << 20,
                                                             320 << 20),
                           steps: int = 50, mode: str = "graph", use_cache: bool = True) -> dict:
    """Time each group size of the ensemble's step schedule (bytes of field data per group,
    0 = all instances in every launch) over `steps`-step calls and apply the fastest."""
    key = f"ensemble_schedule:{sorted(set(ens.shapes))}:{len(ens.shapes)}:{mode}:{steps}"
    if use_cache:
        hit = _load_cache().get(key)
        if hit:
            ens.set_schedule(hit["best"])
            return hit
    results = []
    for gb in candidates:
        ens.set_schedule(int(gb))
        ens.step(steps, mode=mode, timed=True)  # capture / warm
        ms = ens.step(steps, mode=mode, timed=True)
        results.append((ms / steps, int(gb)))
    results.sort()
    best = results[0][1]
    ens.set_schedule(best)
    rec = {"best": best, "us_per_step": results[0][0] * 1e3,
           "table": [[c, t * 1e3] for t, c in results]}
    _store_cache(key, rec)
    return rec


def tune_ensemble_joint(ens, geometries: Optional[Iterable] = None,
                        group_sizes: Iterable[int] = (0, 96 << 20, 160 << 20, 224 << 20,
                                                      320 << 20),
                        probe_group: int = 224 << 20, top: int = 3, steps: int = 100,
                        mode: str = "batched", use_cache: bool = True,
                        concurrency: Iterable[int] = (1, 2)) -> dict:
    """Geometry and step schedule together, in two stages: every geometry under the
    `probe_group` schedule, then the `top` fastest geometries with every group size and every
    number of groups stepping at once (`concurrency`).  The geometry and the group size
    interact (a small group under-fills the GPU when each block marches many planes, unless
    a second group runs beside it), so tuning one with the others fixed can miss the best
    combination.  Tune on a field state like the one to be stepped (the kernels' time depends
    on the state).  best = [geometry, group bytes, groups at once]."""
    gsz = [int(g) for g in group_sizes]
    conc = [int(c) for c in concurrency]
    key = (f"ensemble_joint:{sorted(set(ens.shapes))}:{len(ens.shapes)}:{mode}:{steps}:"
           f"{sorted(gsz)}:{sorted(conc)}")
    if use_cache:
        hit = _load_cache().get(key)
        if hit:
            ens.set_geometry(*hit["best"][0])
            ens.set_schedule(hit["best"][1])
            ens.set_concurrency(hit["best"][2])
            return hit

    def timed(geom, gb, c):
        ens.set_geometry(*geom)
        ens.set_schedule(gb)
        ens.set_concurrency(c)
        ens.step(steps, mode=mode, timed=True)  # warm (and capture, in graph mode)
        return ens.step(steps, mode=mode, timed=True) / steps

    stage1 = sorted((timed(tuple(g), int(probe_group), 1), tuple(g))
                    for g in (geometries or ensemble_candidates()))
    results = []
    for _, g in stage1[:max(1, int(top))]:
        for gb in gsz:
            for c in conc:
                results.append((timed(g, gb, c), g, gb, c))
    results.sort()
    _, g_best, gb_best, c_best = results[0]
    ens.set_geometry(*g_best)
    ens.set_schedule(gb_best)
    ens.set_concurrency(c_best)
    rec = {"best": [list(g_best), gb_best, c_best], "us_per_step": results[0][0] * 1e3,
           "table": [[list(g), gb, c, t * 1e3] for t, g, gb, c in results],
           "geometries": [[list(g), t * 1e3] for t, g in stage1]}
    _store_cache(key, rec)
    return rec
