"""Lattice driver: the reference's main loop around the step, on device-resident fields.

  python -m tip.driver --field real --Nx 64 --Ny 64 --Nz 64 --Nt 200 --T0 7000 --no_hubble
  python -m tip.driver --field complex --Nx 128 --Ny 128 --Nz 128 --Nt 5000 --T0 7350 \\
         --cool_to 6770 --surrogate model.npz

Per step n (simulation/latticeSimeRescale_numba.py:3619-3768; complex
simulation/latticeSimComplex_numba.py:4237-4398):
  * schedule: Hubble (T = T0/a, eta_eff = eta + 3H/mu, inv_a2, inv_a3, a += a H dt/mu,
    :3622-3637), or constant T (--no_hubble), or linear cooling T0 -> --cool_to over Nt steps;
  * V' table: rebuilt on the host when |T - T_last|/T > 1e-4 or the field leaves the table
    range, range = field range +- max(range, 2e4) (:3650-3672; complex uses [0, rho_max + margin],
    :4262-4287).  The field range comes from a device min/max reduction, not a host scan.
    With --surrogate the table is not used: V'(rho, T) comes from the trained network at the
    current T (complex field);
  * integrator (--integrator): rk2 (rk2_step_table), baoab (baoab_step_table) or overdamped
    (overdamped_euler_step_table); "auto" follows the reference: rk2 for the real field, for
    the complex field overdamped when eta_phys > 100 (latticeSimComplex_numba.py:654-666) and
    baoab otherwise (its default, :305); the surrogate path is rk2;
  * noise amplitude: rk2 sqrt(2 eta_eff T dt inv_a3 / (mu^2 dx_phys^3)) (:3714-3718; complex
    :4332-4336 without inv_a3), overdamped sqrt(2 T dt inv_a3 / (eta_eff mu^2 dx_phys^3))
    (:3678; complex :4293), baoab sqrt(T inv_a3 / (mu^2 dx_phys^3)) (:3697; complex :4313),
    with the hash noise stream (noise_seed = step);
  * observables every --report steps (true-vacuum fraction |phi| > esc, string count for the
    complex field) and a NaN/Inf/|phi| > 1e6 * esc guard (:3882-3944);
  * npz snapshots every --steps with the reference's keys (:3996-4018; complex :4632-4654).
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time
from dataclasses import dataclass, field, asdict
from typing import Dict, List, Optional

import numpy as np

from . import potential as P
from .lattice import Lattice

MU = 1000.0
TABLE_MARGIN_FRAC = 1.0  # latticeSimeRescale_numba.py:3604
TABLE_MARGIN_MIN = 20000.0  # latticeSimeRescale_numba.py:3662, latticeSimComplex_numba.py:4277
CHECKPOINT_EVERY = 10    # latticeSimeRescale_numba.py:395


@dataclass
class RunConfig:
    field: str = "real"
    Nx: int = 64
    Ny: int = 64
    Nz: int = 64
    Nt: int = 200
    T0: float = 7000.0
    dx_phys: Optional[float] = None
    dt_phys: Optional[float] = None
    eta_phys: Optional[float] = None
    no_hubble: bool = True
    cool_to: Optional[float] = None
    cool_steps: Optional[int] = None  # length of the linear cooling schedule (default Nt)
    potential_type: Optional[str] = None
    nb: float = 20.0
    nf: float = 20.0
    lam: Optional[float] = None
    zn_order: int = 0
    zn_strength: float = 0.0
    zn_turn_on_T: float = 0.0
    init_rho: float = 0.01
    counterterm: bool = False  # lattice counterterm on the dJ splines (:458-490, :758-795)
    table_source: str = "spline"  # "spline": the reference's V' (dJ spline, no CW);
                                  # "exact": float64 series dJ + Coleman-Weinberg (vprime_full)
    integrator: str = "auto"  # "rk2", "baoab", "overdamped" or "auto" (the reference's choice)
    precision: str = "f32"
    seed: int = 0
    table_size: int = 16384
    esc: float = 1e4
    report: int = 100
    steps: int = 0           # snapshot interval (0 = none)
    outdir: Optional[str] = None
    surrogate: Optional[str] = None
    surrogate_grid: int = 0  # 0: network inline per site; > 0: per-step grid of that many points
    divergence: float = 1e6  # |phi|max threshold of the reference's guard (0 disables)
    guard_every: int = 100000  # latticeSimeRescale_numba.py:3883
    range_every: int = 1     # steps between checks of the field against the table's range (1:
                             # every step, as the reference :3650-3672; > 1: one device min/max
                             # per this many steps, and with constant parameters the steps in
                             # between go to the GPU in one call)
    quiet: bool = False
    progress: int = 0        # log a progress line every this many steps (0 = off)

    def __post_init__(self):
        cplx = self.field == "complex"
        if self.dx_phys is None:
            self.dx_phys = 1e-3
        if self.dt_phys is None:
            self.dt_phys = 2.5e-4 if cplx else 1e-4
        if self.eta_phys is None:
            self.eta_phys = 1000.0 if cplx else self.T0  # complex :641, real eta_phys = T0 :646
        if self.potential_type is None:
            self.potential_type = "V_p" if cplx else "V_correct"
        if self.integrator not in ("auto", "rk2", "baoab", "overdamped"):
            raise ValueError(f"unknown integrator {self.integrator!r}")

    def resolved_integrator(self) -> str:
        if self.integrator != "auto":
            return self.integrator
        if self.field != "complex" or self.surrogate:
            return "rk2"
        return "overdamped" if self.eta_phys > 100.0 else "baoab"


def model_for(cfg: RunConfig) -> P.ModelParams:
    if cfg.field == "complex":
        return P.ModelParams.complex(cfg.potential_type, lam=cfg.lam or 1e-16)
    return P.ModelParams.real(cfg.potential_type, cfg.nb, cfg.nf, lam=cfg.lam or 1e-24)


def initial_fields(cfg: RunConfig):
    """Real: phi = 0.01 N(0,1), pi = 0 (:1259-1260).  Complex: random phase, amplitude
    init_rho |N(0,1)| (latticeSimComplex_numba.py:1271-1279)."""
    rng = np.random.default_rng(cfg.seed)
    shape = (cfg.Nx, cfg.Ny, cfg.Nz)
    if cfg.field == "complex":
        th = 2.0 * np.pi * rng.random(shape)
        r = cfg.init_rho * np.abs(rng.normal(size=shape))
        return {"phi1": r * np.cos(th), "phi2": r * np.sin(th), "pi1": np.zeros(shape),
                "pi2": np.zeros(shape)}
    return {"phi": 0.01 * rng.normal(size=shape), "pi": np.zeros(shape)}


class Run:
    def __init__(self, cfg: RunConfig, log=None):
        self.cfg = cfg
        self.log = log or (lambda m: None if cfg.quiet else print(m, flush=True))
        self.model = model_for(cfg)
        self.cplx = cfg.field == "complex"
        dtype = np.float64 if cfg.precision == "f64" else np.float32
        if self.cplx and dtype == np.float64:
            raise ValueError("the device-resident complex lattice is float32")
        self.lat = Lattice((cfg.Nx, cfg.Ny, cfg.Nz), dtype, complex_field=self.cplx)
        for k, v in initial_fields(cfg).items():
            self.lat.upload(k, v)
        self.dt = MU * cfg.dt_phys
        self.dx = MU * cfg.dx_phys
        self.eta = cfg.eta_phys / MU
        self.hubble = None if cfg.no_hubble else P.HubbleState(cfg.T0, MU, self.eta)
        self.spline = (P.dj_spline(x_cut=P.counterterm_xcut(cfg.dx_phys, cfg.T0))
                       if cfg.counterterm else P.dj_spline())
        self.table_T = None
        self.table_lo = self.table_hi = None
        self.rebuilds = 0
        self.zn_active = 1.0 if (cfg.zn_order > 0 and cfg.zn_turn_on_T <= 0.0) else 0.0
        self.lat.set_zn(cfg.zn_order, cfg.zn_strength, self.zn_active) if self.cplx else None
        self.integrator = cfg.resolved_integrator()
        self.surrogate = None
        if cfg.surrogate and self.integrator != "rk2":
            raise ValueError("the surrogate path integrates with rk2")
        if cfg.surrogate:
            if not self.cplx:
                raise ValueError("the inline surrogate drives the complex field")
            from .surrogate import Surrogate
            self.surrogate = (cfg.surrogate if not isinstance(cfg.surrogate, str)
                              else Surrogate.load(cfg.surrogate))
        self.series: Dict[str, List[float]] = {"step": [], "T": [], "f_true": [], "strings": [],
                                               "rho_max": []}
        self.saves = 0

    # -- schedule ------------------------------------------------------------------------
    def schedule(self, n: int):
        cfg = self.cfg
        if self.hubble is not None:
            s = self.hubble.step(self.dt)
            inv_a3 = s["inv_a3"] if not self.cplx else 1.0
            return s["T"], s["eta_eff"], s["inv_a2"], inv_a3
        if cfg.cool_to is not None:
            T = cfg.T0 + (cfg.cool_to - cfg.T0) * n / max(cfg.cool_steps or cfg.Nt, 1)
        else:
            T = cfg.T0
        return T, self.eta, 1.0, 1.0

    def field_range(self):
        lo, hi = self.lat.minmax()
        if self.cplx:
            return 0.0, hi  # rho max
        return lo, hi

    def maybe_rebuild(self, T: float, n: int = 0) -> None:
        need = self.table_T is None or abs(T - self.table_T) / max(abs(T), 1.0) > 1e-4
        if not need and n % max(self.cfg.range_every, 1) != 0:
            return  # the field's range is checked every range_every steps
        lo, hi = self.field_range()
        if not (math.isfinite(lo) and math.isfinite(hi)):
            raise FloatingPointError("non-finite field")
        if not need:
            need = lo < self.table_lo or hi > self.table_hi
        if not need:
            return
        rng = max(hi - lo, 1.0)
        margin = max(rng * TABLE_MARGIN_FRAC, TABLE_MARGIN_MIN)
        t_lo = max(0.0, lo - margin) if self.cplx else lo - margin
        t_hi = hi + margin
        if self.cfg.table_source == "exact":
            grid = np.linspace(t_lo, t_hi, self.cfg.table_size, dtype=np.float64)
            table = P.vprime_full(grid, T, self.model, include_cw=True)
            tmin, dinv = float(t_lo), 1.0 / ((t_hi - t_lo) / (self.cfg.table_size - 1))
            self.lat.set_table(table, tmin, dinv, odd=False)
        else:
            # the reference's table producer (Vprime_scalar over the grid) on the device,
            # straight into the lattice's table buffer
            self.lat.build_table(T, t_lo, t_hi, self.cfg.table_size, self.model, self.spline)
        self.table_T, self.table_lo, self.table_hi = T, t_lo, t_hi
        self.rebuilds += 1

    def noise_scale(self, T, eta_eff, inv_a3) -> float:
        dx3 = MU ** 2 * self.cfg.dx_phys ** 3
        if self.integrator == "overdamped":
            return math.sqrt(2.0 * T * self.dt * inv_a3 / (eta_eff * dx3))
        if self.integrator == "baoab":
            return math.sqrt(T * inv_a3 / dx3)
        return math.sqrt(2.0 * eta_eff * T * self.dt * inv_a3 / dx3)

    # -- observables ---------------------------------------------------------------------
    def guard(self, n: int, T: float) -> None:
        """NaN / Inf / divergence guard (latticeSimeRescale_numba.py:3882-3944): dump the state
        and stop."""
        lo, hi = self.lat.minmax()
        bad = not (math.isfinite(lo) and math.isfinite(hi))
        if not bad and self.cfg.divergence > 0 and n % self.cfg.guard_every == 0:
            bad = max(abs(lo), abs(hi)) > self.cfg.divergence
        if bad:
            self.dump_debug(n, T)
            raise FloatingPointError(f"field diverged at step {n} (min {lo:.3e}, max {hi:.3e})")

    def observe(self, n: int, T: float) -> None:
        f = self.lat.vacuum_fraction(self.cfg.esc)
        strings = self.lat.winding_counts()[0] if self.cplx else 0
        lo, hi = self.lat.minmax()
        self.series["step"].append(n)
        self.series["T"].append(T)
        self.series["f_true"].append(f)
        self.series["strings"].append(strings)
        self.series["rho_max"].append(hi)

    def dump_debug(self, n: int, T: float) -> None:
        if not self.cfg.outdir:
            return
        os.makedirs(self.cfg.outdir, exist_ok=True)
        names = ["phi1", "phi2", "pi1", "pi2"] if self.cplx else ["phi", "pi"]
        np.savez(os.path.join(self.cfg.outdir, f"state_step_{n:010d}_NaN_debug.npz"),
                 step=n, temperature=T, **{k: self.lat.download(k) for k in names})

    def snapshot(self, n: int, T: float) -> None:
        cfg = self.cfg
        os.makedirs(cfg.outdir, exist_ok=True)
        ckpt = self.saves % CHECKPOINT_EVERY == 0
        self.saves += 1
        d = dict(step=n, time=n * self.dt, temperature=T)
        if self.cplx:
            sp = self.lat.save_point()  # rho, theta, winding buffer and rho range on the device
            d.update(phi1=self.lat.download("phi1"), phi2=self.lat.download("phi2"),
                     rho=sp["rho"], theta=sp["theta"], winding=sp["winding"],
                     rho_min=sp["rho_min"], rho_max=sp["rho_max"],
                     zn_active=float(self.zn_active))
            if ckpt and self.integrator != "overdamped":
                d.update(pi1=self.lat.download("pi1"), pi2=self.lat.download("pi2"))
        else:
            phi = self.lat.download("phi")
            d.update(phi=phi, phi_min=float(phi.min()), phi_max=float(phi.max()))
            if ckpt and self.integrator != "overdamped":
                d["pi"] = self.lat.download("pi")
        if self.hubble is not None:
            d.update(scale_factor=self.hubble.a)
        np.savez_compressed(os.path.join(cfg.outdir, f"state_step_{n:010d}.npz"), **d)

    # -- main loop -----------------------------------------------------------------------
    def _chunk(self, n: int) -> int:
        """Steps until the next report / guard / snapshot / progress / range-check boundary (the
        surrogate path: the temperature is linear in n inside a chunk; the table path with
        constant parameters: nothing changes inside a chunk; either way one C call runs it)."""
        cfg = self.cfg
        k = cfg.Nt - n
        for every in (cfg.report, cfg.guard_every, cfg.steps if cfg.outdir else 0, cfg.progress,
                      cfg.range_every if self.surrogate is None else 0):
            if every:
                k = min(k, every - n % every)
        return max(k, 1)

    def run(self) -> Dict[str, list]:
        cfg = self.cfg
        t0 = time.perf_counter()
        n = 0
        while n < cfg.Nt:
            T, eta_eff, inv_a2, inv_a3 = self.schedule(n)
            if self.cplx and cfg.zn_order > 0 and cfg.zn_turn_on_T > 0 and T < cfg.zn_turn_on_T \
                    and self.zn_active < 0.5:
                self.zn_active = 1.0
                self.lat.set_zn(cfg.zn_order, cfg.zn_strength, 1.0)
            ns = self.noise_scale(T, eta_eff, inv_a3)
            kw = dict(dt=self.dt, dx=self.dx, eta_eff=eta_eff, mu=MU, inv_a2=inv_a2,
                      noise_scale=ns, seed=cfg.seed, step0=n)
            if self.surrogate is not None and self.hubble is None and not (
                    cfg.zn_order > 0 and cfg.zn_turn_on_T > 0 and self.zn_active < 0.5):
                k = self._chunk(n)
                dT = (0.0 if cfg.cool_to is None
                      else (cfg.cool_to - cfg.T0) / max(cfg.cool_steps or cfg.Nt, 1))
                self.lat.step_mlp(k, self.surrogate, T_start=T, dT_per_step=dT,
                                  grid_points=cfg.surrogate_grid, **kw)
                T = T + dT * (k - 1)
            elif self.surrogate is not None:
                k = 1
                self.lat.step_mlp(1, self.surrogate, T_start=T, grid_points=cfg.surrogate_grid,
                                  **kw)
            else:
                # constant parameters (no expansion, no cooling, no Z_N switch ahead): the steps
                # up to the next check, report or snapshot go to the GPU in one call
                const = (self.hubble is None and cfg.cool_to is None and
                         not (cfg.zn_order > 0 and cfg.zn_turn_on_T > 0 and self.zn_active < 0.5))
                k = self._chunk(n) if const else 1
                self.maybe_rebuild(T, n)
                if self.integrator == "baoab":
                    kw["gamma"] = kw.pop("eta_eff")
                    self.lat.step_baoab(k, **kw)
                elif self.integrator == "overdamped":
                    self.lat.step_overdamped(k, **kw)
                else:
                    self.lat.step(k, **kw)
            n += k
            T_now = T  # temperature of the last step taken
            if cfg.report and n % cfg.report == 0:
                self.observe(n, T_now)
            if n % min(cfg.guard_every, cfg.report or cfg.guard_every) == 0:
                self.guard(n, T_now)
            if cfg.progress and n % cfg.progress == 0:
                el = time.perf_counter() - t0
                f = self.series["f_true"][-1] if self.series["f_true"] else float("nan")
                print(f"[run] step {n}/{cfg.Nt} T={T_now:.2f} {1e3 * el / n:.3f} ms/step "
                      f"f_true={f:.4f}", flush=True)
            if cfg.steps and cfg.outdir and n % cfg.steps == 0:
                self.snapshot(n, T_now)
                el = time.perf_counter() - t0
                f = self.series["f_true"][-1] if self.series["f_true"] else float("nan")
                self.log(f"Step {n}/{cfg.Nt} | T={T_now:.1f} | {n / el:.1f} steps/s | "
                         f"{1e3 * el / n:.2f} ms/step | f_true={f:.4f}")
        self.elapsed = time.perf_counter() - t0
        return self.series

    def close(self):
        self.lat.close()


def nucleation_ensemble(steps: int = 20000, report: int = 500, esc: float = 1e4,
                        frac: float = 0.5, mode: str = "graph", log=None, geometry=None,
                        group_bytes: Optional[int] = None, concurrency: Optional[int] = None):
    """The nucleation ensemble: 64 realisations (32 x 64^3, 24 x 96^3, 8 x 128^3), seeds 1..64,
    T0 ~ U[6770, 7350], per-instance 8192-point tables; returns each instance's nucleation step
    (first report step with true-vacuum fraction > frac; -1 if none) and the fraction series.
    geometry / group_bytes / concurrency: the ensemble's launch settings (Ensemble.set_geometry,
    set_schedule, set_concurrency; None keeps the defaults), e.g. an autotuner's choice."""
    from .ensemble import Ensemble, nucleation_ensemble_shapes

    shapes = nucleation_ensemble_shapes()
    rng = np.random.default_rng(2024)
    T0s = rng.uniform(6770.0, 7350.0, size=len(shapes))
    p = P.ModelParams.real()
    ens = Ensemble(shapes, seeds=list(range(1, len(shapes) + 1)))
    dt, dx = MU * 2.5e-4, MU * 0.002
    for i, T0 in enumerate(T0s):
        vt = P.thermal_vev(float(T0), p)
        t, tmin, dinv = P.build_vprime_table(float(T0), 0.0, 1.5 * vt, 8192, p)
        ens.set_table(i, t, tmin, dinv, odd=True)
        eta = T0 / MU
        ens.set_params(i, dt=dt, dx=dx, eta_eff=eta, mu=MU,
                       noise_scale=P.noise_scale(eta, float(T0), dt, MU, 0.002))
    r2 = np.random.default_rng(0)
    for i, s in enumerate(shapes):
        ens.upload(i, "phi", 0.01 * r2.normal(size=s))
        ens.upload(i, "pi", np.zeros(s))
    if geometry is not None:
        ens.set_geometry(*geometry)
    if group_bytes is not None:
        ens.set_schedule(int(group_bytes))
    if concurrency is not None:
        ens.set_concurrency(int(concurrency))
    nuc = np.full(len(shapes), -1, dtype=np.int64)
    series = []
    t0 = time.perf_counter()
    done = 0
    while done < steps:
        k = min(report, steps - done)
        ens.step(k, mode=mode)
        done += k
        f = ens.vacuum_fractions(esc)
        series.append(f)
        newly = (nuc < 0) & (f > frac)
        nuc[newly] = done
        if log:
            log(f"step {done}: nucleated {int((nuc >= 0).sum())}/{len(shapes)}")
    el = time.perf_counter() - t0
    ens.close()
    return {"T0": T0s, "nucleation_step": nuc, "fractions": np.array(series), "elapsed_s": el,
            "site_updates_per_s": float(sum(int(np.prod(s)) for s in shapes)) * steps / el}


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    for f in RunConfig.__dataclass_fields__.values():
        name = "--" + f.name
        if f.type in ("bool", bool):
            ap.add_argument(name, action="store_true")
        else:
            typ = {"int": int, "float": float, "str": str}.get(str(f.type).replace("Optional[", "").rstrip("]"), str)
            ap.add_argument(name, type=typ, default=None)
    ap.add_argument("--hubble", action="store_true", help="enable Hubble expansion")
    ap.add_argument("--json", default=None, help="write the observable series here")
    a = vars(ap.parse_args(argv))
    hub = a.pop("hubble")
    js = a.pop("json")
    kw = {k: v for k, v in a.items() if v is not None and v is not False}
    cfg = RunConfig(**kw)
    if hub:
        cfg.no_hubble = False
    run = Run(cfg)
    try:
        series = run.run()
    finally:
        run.close()
    out = {"config": asdict(cfg), "series": series, "elapsed_s": run.elapsed,
           "table_rebuilds": run.rebuilds}
    if js:
        with open(js, "w") as fh:
            json.dump(out, fh)
    print(json.dumps({"elapsed_s": run.elapsed, "steps": cfg.Nt,
                      "ms_per_step": 1e3 * run.elapsed / max(cfg.Nt, 1),
                      "final_f_true": series["f_true"][-1] if series["f_true"] else None,
                      "table_rebuilds": run.rebuilds}))
    return 0


if __name__ == "__main__":
    sys.exit(main())
