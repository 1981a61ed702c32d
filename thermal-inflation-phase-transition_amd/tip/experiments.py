"""Experiments on top of the driver: the ensemble nucleation run (64 realisations of
64^3-128^3) and the surrogate-versus-table comparison under continuous cooling (7350 -> 6770 K
over 10^5 steps).

  python -m tip.experiments surrogate-vs-table --n 128 --steps 3000 --seeds 6 --report 100 \
         --cool-steps 100000 --exact-table
  python -m tip.experiments nucleation --steps 20000
"""
from __future__ import annotations

import argparse
import json
import sys
import time
from typing import Dict, Sequence

import numpy as np

from .driver import Run, RunConfig, nucleation_ensemble


def surrogate_vs_table(surrogate, n: int = 64, steps: int = 4000, T0: float = 7350.0,
                       T1: float = 6770.0, seeds: Sequence[int] = (0, 1, 2, 3),
                       report: int = 250, esc: float = 1e4, cool_steps: int = None,
                       methods: Sequence[str] = ("table", "surrogate"), log=None) -> Dict:
    """Complex field, linear cooling T0 -> T1 over `cool_steps` steps (default `steps`; the
    full cooling run is 10^5 steps, of which the first `steps` are run), an ensemble of seeds
    run with each method: "table" -- the reference's table method (host-built 16384-point
    table of the reference's V' on the field range, rebuilt when |dT|/T > 1e-4 or the field
    leaves the range); "surrogate" -- the inline network at the exact T(t); "table_exact"
    (optional) -- the table method fed the float64 potential the network was trained on
    (series J + Coleman-Weinberg), which separates the network's error from the reference's
    dJ-spline error.  Returns per-report-step ensemble means and standard deviations of the
    true-vacuum fraction and the string count, and whether table and surrogate agree within
    the combined 1 sigma, sqrt(sd_t^2 + sd_s^2), at every report step."""
    out = {}
    for method in methods:
        F, S, el, rebuilds = [], [], 0.0, 0
        for sd in seeds:
            cfg = RunConfig(field="complex", integrator="rk2", Nx=n, Ny=n, Nz=n, Nt=steps,
                            T0=T0, cool_to=T1,
                            cool_steps=cool_steps, seed=sd, report=report, esc=esc, quiet=True,
                            surrogate=surrogate if method == "surrogate" else None,
                            table_source="exact" if method == "table_exact" else "spline",
                            divergence=0.0)
            run = Run(cfg)
            try:
                ser = run.run()
            finally:
                run.close()
            F.append(ser["f_true"])
            S.append(ser["strings"])
            el += run.elapsed
            rebuilds += run.rebuilds
            if log:
                log(f"{method} seed {sd}: {run.elapsed:.1f}s, final f_true {ser['f_true'][-1]:.4f}, "
                    f"strings {ser['strings'][-1]}")
        F, S = np.array(F, dtype=np.float64), np.array(S, dtype=np.float64)
        out[method] = {"f_mean": F.mean(0), "f_sd": F.std(0, ddof=1) if len(seeds) > 1 else 0 * F[0],
                       "s_mean": S.mean(0), "s_sd": S.std(0, ddof=1) if len(seeds) > 1 else 0 * S[0],
                       "elapsed_s": el, "ms_per_step": 1e3 * el / (steps * len(seeds)),
                       "table_rebuilds": rebuilds}
        out["steps_axis"] = list(range(report, steps + 1, report))
    def within(m1, s1, m2, s2):
        tol = np.sqrt(s1 ** 2 + s2 ** 2)
        return np.abs(m1 - m2) <= np.maximum(tol, 1e-12)

    for ref in ("table", "table_exact"):
        if ref in out and "surrogate" in out:
            t, s = out[ref], out["surrogate"]
            sfx = "" if ref == "table" else "_exact"
            out["f_within_1sigma" + sfx] = within(t["f_mean"], t["f_sd"], s["f_mean"], s["f_sd"])
            out["s_within_1sigma" + sfx] = within(t["s_mean"], t["s_sd"], s["s_mean"], s["s_sd"])
    return out


def _jsonable(o):
    if isinstance(o, dict):
        return {k: _jsonable(v) for k, v in o.items()}
    if isinstance(o, (np.ndarray, list, tuple)):
        return [_jsonable(v) for v in o]
    if isinstance(o, (np.floating, np.integer, np.bool_)):
        return o.item()
    return o


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="what", required=True)
    a1 = sub.add_parser("surrogate-vs-table")
    a1.add_argument("--surrogate", default=None,
                    help="trained npz (default: the shipped cooling-run weights)")
    a1.add_argument("--train", action="store_true", help="train a fresh network first")
    a1.add_argument("--n", type=int, default=64)
    a1.add_argument("--steps", type=int, default=4000)
    a1.add_argument("--seeds", type=int, default=4)
    a1.add_argument("--report", type=int, default=250)
    a1.add_argument("--cool-steps", type=int, default=None,
                    help="length of the cooling schedule (the full run: 100000)")
    a1.add_argument("--exact-table", action="store_true",
                    help="also run the table method on the float64 potential")
    a2 = sub.add_parser("nucleation")
    a2.add_argument("--steps", type=int, default=20000)
    a2.add_argument("--report", type=int, default=500)
    ap.add_argument("--json", default=None)
    a = ap.parse_args(argv)
    log = lambda m: print(m, file=sys.stderr, flush=True)
    if a.what == "surrogate-vs-table":
        from .surrogate import Surrogate, train_cooling_surrogate
        sur = (Surrogate.load(a.surrogate) if a.surrogate
               else train_cooling_surrogate(log=log)[0] if a.train else Surrogate.load())
        res = surrogate_vs_table(sur, n=a.n, steps=a.steps, seeds=range(a.seeds),
                                 report=a.report, cool_steps=a.cool_steps,
                                 methods=("table", "surrogate") + (("table_exact",)
                                                                   if a.exact_table else ()),
                                 log=log)
    else:
        res = nucleation_ensemble(steps=a.steps, report=a.report, log=log)
    res = _jsonable(res)
    if a.json:
        with open(a.json, "w") as fh:
            json.dump(res, fh)
    print(json.dumps({k: v for k, v in res.items() if k not in ("fractions",)})[:4000])
    return 0


if __name__ == "__main__":
    sys.exit(main())
