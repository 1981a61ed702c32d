"""Benchmark of the fused RK2 Langevin lattice step on one MI355X.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload real256] [--no-cpu]

Workload (default "real256"): real scalar field on a 256^3 periodic lattice, float32,
V' table of 8192 float64 points on [0, 1.5 phi_true] (odd) built at T0 = 7000 GeV from the
tree + thermal potential (lambda = 1e-24, V_correct, nb = nf = 20, y = 1.09, g = 1.05),
dx = 2.0 (dx_phys = 0.002), dt = 0.25 (dt_phys = 2.5e-4), eta = T0/mu = 7, hash noise with
the fluctuation-dissipation amplitude; phi0 = 0.01 N(0, 1), pi0 = 0.  One step = one RK2
step over the whole lattice = one fused kernel launch (state resident in HBM, ping-pong).

Prints one JSON line: site-updates/s (value), HBM GB/s at the algorithmic 16 B/site, the
roofline of the fused kernel (HIP-event time on the launch stream) and the CPU baseline (the
oracle's float32 two-pass restatement of rk2_step_table with its separate hash-noise pass,
on all host cores, over a bounded sample of steps).
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
PKG = ROOT / "thermal-inflation-phase-transition_amd"
for _p in (str(ROOT), str(PKG)):
    if _p not in sys.path:
        sys.path.insert(0, _p)

HBM_PEAK_GBPS = 8000.0
REF_CPU_SITE_UPDATES = 2.21e8  # reference log, rk2_step_table 256^3 float32, 11 threads


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def real_workload(n: int = 256, npts: int = 8192):
    from tip import potential as P

    p = P.ModelParams.real()
    T0, mu, dx_phys, dt_phys = 7000.0, 1000.0, 0.002, 2.5e-4
    vt = P.thermal_vev(T0, p)
    table, tmin, dinv = P.build_vprime_table(T0, 0.0, 1.5 * vt, npts, p)
    dt, dx, eta = mu * dt_phys, mu * dx_phys, T0 / mu
    prm = dict(dt=dt, dx=dx, eta_eff=eta, mu=mu, inv_a2=1.0,
               noise_scale=P.noise_scale(eta, T0, dt, mu, dx_phys))
    return table, tmin, dinv, prm


def profile_record(name: str) -> dict:
    """The `key: value` lines of a committed rocprofv3 summary under profiles/ (HBM bytes per
    launch from FETCH_SIZE / WRITE_SIZE, the kernel-trace average of the same workload)."""
    f = ROOT / "profiles" / name
    out = {}
    if not f.exists():
        return out
    for line in f.read_text().splitlines():
        if ":" in line and not line.startswith(("#", "|", " ")):
            k, v = line.split(":", 1)
            try:
                out[k.strip()] = float(v.strip())
            except ValueError:
                pass
    return out


def host_description() -> dict:
    """CPU model, SMT state and core counts of the host the CPU baseline runs on."""
    model = None
    ident = {}
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            key = line.split(":", 1)[0].strip()
            if key in ("model name", "cpu family", "model", "stepping", "cpu MHz") and \
                    key not in ident:
                ident[key] = line.split(":", 1)[1].strip()
            if not line.strip() and ident:
                break  # first processor's block only
        model = ident.get("model name")
    except OSError:
        pass
    smt = None
    try:
        smt = Path("/sys/devices/system/cpu/smt/active").read_text().strip() == "1"
    except OSError:
        pass
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count()
    return {"cpu_model": model, "cpu_family": ident.get("cpu family"),
            "cpu_model_number": ident.get("model"), "cpu_stepping": ident.get("stepping"),
            "cpu_mhz": ident.get("cpu MHz"), "smt_active": smt, "nproc": avail,
            "os_cpu_count": os.cpu_count()}


def cpu_baseline(shape, table, tmin, dinv, prm, budget_s: float = 12.0):
    """Oracle float32 restatement of rk2_step_table (2 passes) + hash noise pass, timed."""
    from oracle import oracle as O

    rng = np.random.default_rng(1)
    phi = (0.01 * rng.normal(size=shape)).astype(np.float32)
    pi = np.zeros(shape, np.float32)
    pt, qt = np.empty_like(phi), np.empty_like(pi)
    noise = np.empty(shape, np.float32)

    def one(n):
        O.lib().orc_noise_field_f32(noise, None, *shape, O.noise_seed(n), 73856093,
                                    prm["noise_scale"])
        O.rk2_step_table(phi, pi, prm["dt"], prm["dx"], prm["eta_eff"], prm["mu"], noise, pt, qt,
                         prm["inv_a2"], table, tmin, dinv, len(table), odd=True)

    one(0)  # warm-up (page faults, thread pool)
    steps, t0 = 0, time.perf_counter()
    while True:
        one(steps + 1)
        steps += 1
        el = time.perf_counter() - t0
        if el >= budget_s or steps >= 200:
            break
    sites = float(np.prod(shape))
    return dict(value=sites * steps / el, unit="site-updates/s", cores=O.num_threads(),
                kind="port",
                sample=(f"{steps} steps of {shape[0]}x{shape[1]}x{shape[2]} real float32: oracle "
                        f"rk2_step_table (2 OpenMP passes, float32 fields, float64 arithmetic) "
                        f"+ hash noise pass, {el:.1f} s"),
                ms_per_step=1e3 * el / steps)


def cpu_baseline_ensemble(budget_s: float = 6.0) -> dict:
    """The nucleation ensemble (64 realisations, 32 x 64^3 + 24 x 96^3 + 8 x 128^3, per-instance
    tables) stepped as a sequential loop on the host cores: one oracle float32 rk2_step_table
    (+ its hash noise pass) per instance per step."""
    from oracle import oracle as O

    shapes, seeds, _, inst = ensemble_workload()
    rng = np.random.default_rng(0)
    fields = []
    for s in shapes:
        phi = (0.01 * rng.normal(size=s)).astype(np.float32)
        fields.append((phi, np.zeros(s, np.float32), np.empty(s, np.float32),
                       np.empty(s, np.float32), np.empty(s, np.float32)))

    def one(n):
        for (phi, pi, pt, qt, noise), sd, (t, tmin, dinv, prm) in zip(fields, seeds, inst):
            O.lib().orc_noise_field_f32(noise, None, *phi.shape, O.noise_seed(n, sd), 73856093,
                                        prm["noise_scale"])
            O.rk2_step_table(phi, pi, prm["dt"], prm["dx"], prm["eta_eff"], prm["mu"], noise, pt,
                             qt, prm["inv_a2"], t, tmin, dinv, len(t), odd=True)

    one(0)
    steps, t0 = 0, time.perf_counter()
    while True:
        one(steps + 1)
        steps += 1
        el = time.perf_counter() - t0
        if el >= budget_s or steps >= 50:
            break
    sites = float(sum(int(np.prod(s)) for s in shapes))
    return dict(value=sites * steps / el, unit="site-updates/s", cores=O.num_threads(),
                kind="port", ms_per_step=1e3 * el / steps,
                sample=(f"{steps} ensemble steps (64 realisations, {int(sites)} sites) as a "
                        f"sequential per-instance loop of the oracle's float32 rk2_step_table "
                        f"+ hash noise pass, {el:.1f} s"))


def cpu_baseline_integrators(table, tmin, dinv, prm, n: int = 128, reps: int = 3) -> dict:
    """The oracle's BAOAB (open + close) and overdamped steps (float64, OpenMP) at n^3 on the
    host cores, for the BAOAB / overdamped lines of the extras."""
    from oracle import oracle as O

    shape = (n, n, n)
    rng = np.random.default_rng(3)
    phi = 0.01 * rng.normal(size=shape)
    pi = np.zeros(shape)
    out = {"cores": O.num_threads(), "kind": "port", "unit": "site-updates/s"}
    for kind in ("baoab", "overdamped"):
        def one(k):
            if kind == "baoab":
                return O.baoab_step(phi, pi, prm["dt"], prm["dx"], prm["eta_eff"], prm["mu"],
                                    1.0, table, tmin, dinv, odd=True, noise_seed=k,
                                    noise_scale=30.0)
            return O.langevin(O.OVERDAMPED, phi, None, prm["dt"], prm["dx"], prm["eta_eff"],
                              prm["mu"], 1.0, table, tmin, dinv, odd=True, noise_seed=k,
                              noise_scale=30.0)
        one(0)
        t0 = time.perf_counter()
        for k in range(reps):
            one(k + 1)
        el = (time.perf_counter() - t0) / reps
        out[kind] = {"value": n ** 3 / el, "ms_per_step": 1e3 * el,
                     "sample": f"{reps} steps of {n}^3 real float64 (oracle restatement)"}
        log(f"cpu {kind} {n}^3: {1e3 * el:.1f} ms/step")
    return out


def ensemble_workload(npts: int = 8192, seed: int = 2024):
    """64 realisations: 32 x 64^3, 24 x 96^3, 8 x 128^3; seeds 1..64; T0 ~ U[6770, 7350] GeV;
    per-instance 8192-point float64 V' tables on [0, 1.5 phi_true(T0)] (odd)."""
    from tip import potential as P
    from tip.ensemble import nucleation_ensemble_shapes

    shapes = nucleation_ensemble_shapes()
    rng = np.random.default_rng(seed)
    T0s = rng.uniform(6770.0, 7350.0, size=len(shapes))
    p = P.ModelParams.real()
    mu, dx_phys, dt_phys = 1000.0, 0.002, 2.5e-4
    inst = []
    for i, T0 in enumerate(T0s):
        vt = P.thermal_vev(float(T0), p)
        t, tmin, dinv = P.build_vprime_table(float(T0), 0.0, 1.5 * vt, npts, p)
        eta = T0 / mu
        prm = dict(dt=mu * dt_phys, dx=mu * dx_phys, eta_eff=eta, mu=mu, inv_a2=1.0,
                   noise_scale=P.noise_scale(eta, float(T0), mu * dt_phys, mu, dx_phys))
        inst.append((t, tmin, dinv, prm))
    return shapes, list(range(1, len(shapes) + 1)), T0s, inst


def bench_ensemble(steps: int = 200, modes=("graph", "batched", "sequential")):
    from tip import autotune
    from tip.ensemble import Ensemble

    shapes, seeds, T0s, inst = ensemble_workload()
    ens = Ensemble(shapes, seeds)
    for i, (t, tmin, dinv, prm) in enumerate(inst):
        ens.set_table(i, t, tmin, dinv, odd=True)
        ens.set_params(i, **prm)
    rng = np.random.default_rng(0)
    for i, s in enumerate(shapes):
        ens.upload(i, "phi", 0.01 * rng.normal(size=s))
        ens.upload(i, "pi", np.zeros(s))
    # every mode is timed from the same state, past the initial roll to the true vacuum; the
    # launch geometry and the step schedule are tuned together on that state
    ens.step(300, mode="batched")
    state = [(ens.download(i, "phi"), ens.download(i, "pi")) for i in range(len(shapes))]
    tune = autotune.tune_ensemble_joint(ens, steps=100, use_cache=False)
    log(f"ensemble autotune best geometry {tune['best'][0]}, {tune['best'][1]} bytes/group, "
        f"{tune['best'][2]} group(s) at once: "
        f"{tune['us_per_step']:.1f} us/step ({tune['table']})")
    sites = float(sum(int(np.prod(s)) for s in shapes))
    out = {"instances": len(shapes), "sites_per_step": sites, "autotune_best": tune["best"][0],
           "group_bytes": tune["best"][1], "groups_at_once": tune["best"][2]}
    def restore():
        for i, (phi, pi) in enumerate(state):
            ens.upload(i, "phi", phi)
            ens.upload(i, "pi", pi)

    for mode in modes:
        # a warm-up call of the timed length (graph mode captures its chunk graphs on first
        # use, outside the timed call), then from the common state 10 steps and the timed call
        restore()
        ens.step(steps, mode=mode, timed=True)
        restore()
        ens.step(10, mode=mode, timed=True)
        t0 = time.perf_counter()
        ms = ens.step(steps, mode=mode, timed=True)
        wall = time.perf_counter() - t0
        out[mode] = {"site_updates_per_s": sites * steps / wall, "ms_per_step": 1e3 * wall / steps,
                     "event_ms_per_step": ms / steps}
        log(f"ensemble {mode}: {1e3 * wall / steps:.3f} ms/step")
    if "sequential" in out:
        seq = out["sequential"]["site_updates_per_s"]
        ratios = []
        for mode in ("graph", "batched"):
            if mode in out:
                out[f"{mode}_vs_sequential"] = out[mode]["site_updates_per_s"] / seq
                ratios.append(out[f"{mode}_vs_sequential"])
        # the grouped ensemble's rate: the faster of its two launch forms (graph replay or
        # direct launches of the same grouped kernels) against the plain per-instance loop
        if ratios:
            out["grouped_vs_sequential"] = max(ratios)
    for mode in ("graph", "batched"):
        if mode in out:
            out[f"hbm_gbps_{mode}"] = 16.0 * out[mode]["site_updates_per_s"] / 1e9
    ens.close()
    return out


def bench_config_runs(ensemble_settings=None) -> dict:
    """The 256^3 workloads end to end at their own lengths, observables included: the real
    field for 10^4 RK2 steps in float32 and in the float64 drop-in precision with the
    true-vacuum fraction every 500 steps (the 8192-point table on [0, 1.5 phi_true] at 7000
    GeV), and the complex field for 5 x 10^4 steps at 6770 GeV with the plaquette windings and
    string count every 500 steps.  Rates count the observables' time too."""
    from tip import potential as P
    from tip.lattice import Lattice

    n = 256
    shape = (n, n, n)
    sites = float(n ** 3)
    out = {}
    table, tmin, dinv, prm = real_workload(n)
    kw = dict(dt=prm["dt"], dx=prm["dx"], eta_eff=prm["eta_eff"], mu=prm["mu"],
              noise_scale=prm["noise_scale"])
    for dtype, name in ((np.float32, "real256_f32_1e4"), (np.float64, "real256_f64_1e4")):
        lat = Lattice(shape, dtype)
        lat.upload("phi", 0.01 * np.random.default_rng(0).normal(size=shape))
        lat.upload("pi", np.zeros(shape))
        lat.set_table(table, tmin, dinv, odd=True)
        total, report = 10000, 500
        f_true = []
        lat.step(1, **kw)  # first launch (module load) outside the clock
        lat.upload("phi", 0.01 * np.random.default_rng(0).normal(size=shape))
        lat.upload("pi", np.zeros(shape))
        t0 = time.perf_counter()
        for k in range(total // report):
            lat.step(report, step0=k * report, **kw)
            f_true.append(lat.vacuum_fraction(1e4))
        wall = time.perf_counter() - t0
        lat.close()
        first = next((report * (i + 1) for i, f in enumerate(f_true) if f > 0.5), None)
        out[name] = {"steps": total, "wall_s": wall, "ms_per_step": 1e3 * wall / total,
                     "site_updates_per_s": sites * total / wall,
                     "hbm_gbps_algorithmic": (16.0 if dtype == np.float32 else 32.0) * sites *
                     total / wall / 1e9,
                     "f_true_every_500": f_true, "first_step_f_true_above_half": first}
        log(f"config run {name}: {1e3 * wall / total:.4f} ms/step over {total} steps with the "
            f"vacuum fraction every {report}; f_true > 0.5 from step {first}")
    mu, T0 = 1000.0, 6770.0
    pc = P.ModelParams.complex()
    vt = P.thermal_vev(T0, pc)
    t, tmin, dinv = P.build_vprime_table(T0, 0.0, 1.5 * vt, 8192, pc)
    dt, dx, eta = mu * 2.5e-4, mu * 1e-3, 1000.0 / mu
    kwc = dict(dt=dt, dx=dx, eta_eff=eta, mu=mu, inv_a2=1.0,
               noise_scale=P.noise_scale(eta, T0, dt, mu, 1e-3))
    rng = np.random.default_rng(2)
    th = 2 * np.pi * rng.random(shape)
    r = 0.01 * np.abs(rng.normal(size=shape))
    lat = Lattice(shape, np.float32, complex_field=True)
    lat.set_table(t, tmin, dinv, odd=False)
    lat.step(1, **kwc)
    lat.upload("phi1", r * np.cos(th))
    lat.upload("phi2", r * np.sin(th))
    lat.upload("pi1", np.zeros(shape))
    lat.upload("pi2", np.zeros(shape))
    total, report = 50000, 500
    strings = []
    t0 = time.perf_counter()
    for k in range(total // report):
        lat.step(report, step0=k * report, **kwc)
        strings.append(int(lat.winding_counts()[0]))
    wall = time.perf_counter() - t0
    lat.close()
    out["complex256_5e4"] = {"steps": total, "wall_s": wall, "ms_per_step": 1e3 * wall / total,
                             "site_updates_per_s": sites * total / wall,
                             "hbm_gbps_algorithmic": 32.0 * sites * total / wall / 1e9,
                             "strings_every_500": strings}
    log(f"config run complex256_5e4: {1e3 * wall / total:.4f} ms/step over {total} steps with "
        f"the windings every {report}; strings {strings[0]} -> {strings[-1]}")
    # the nucleation ensemble: 2 x 10^4 steps, per-instance vacuum fractions every 500 steps,
    # graph replay with the launch settings the ensemble bench tuned (if it ran)
    from tip.driver import nucleation_ensemble
    geo, gb, conc = ensemble_settings or (None, None, None)
    r = nucleation_ensemble(steps=20000, report=500, mode="graph", geometry=geo,
                            group_bytes=gb, concurrency=conc)
    nuc = [int(v) for v in r["nucleation_step"]]
    out["nucleation_ensemble_2e4"] = {
        "steps": 20000, "wall_s": r["elapsed_s"], "ms_per_step": 1e3 * r["elapsed_s"] / 20000,
        "site_updates_per_s": r["site_updates_per_s"],
        "hbm_gbps_algorithmic": 16.0 * r["site_updates_per_s"] / 1e9,
        "nucleation_step": nuc, "instances_nucleated": sum(v >= 0 for v in nuc),
        "launch_settings": [list(geo) if geo is not None else None, gb, conc]}
    log(f"config run nucleation_ensemble_2e4: {1e3 * r['elapsed_s'] / 20000:.4f} ms/step, "
        f"{sum(v >= 0 for v in nuc)}/64 nucleated by step {max(nuc)}")
    return out


def bench_extras(steps: int = 200) -> dict:
    """The other forms on one GPU: complex 256^3 float32 step and winding, the float64 drop-in
    real 256^3 step, the complex 512^3 step with the inline MLP surrogate (the 7350 -> 6770 K
    cooling workload's lattice; throughput does not depend on the weights, untrained ones are
    used), the BAOAB and overdamped steps, the component labelling and the V' table
    producer."""
    from tip import potential as P
    from tip import surrogate as S
    from tip.lattice import Lattice

    out = {}
    mu, T0 = 1000.0, 6770.0
    n = 256
    pc = P.ModelParams.complex()
    vt = P.thermal_vev(T0, pc)
    t, tmin, dinv = P.build_vprime_table(T0, 0.0, 1.5 * vt, 8192, pc)
    dt, dx, eta = mu * 2.5e-4, mu * 1e-3, 1000.0 / mu
    kw = dict(dt=dt, dx=dx, eta_eff=eta, mu=mu, inv_a2=1.0,
              noise_scale=P.noise_scale(eta, T0, dt, mu, 1e-3))
    shape = (n, n, n)
    rng = np.random.default_rng(2)
    th = 2 * np.pi * rng.random(shape)
    r = 0.01 * np.abs(rng.normal(size=shape))
    lat = Lattice(shape, np.float32, complex_field=True)
    lat.upload("phi1", r * np.cos(th))
    lat.upload("phi2", r * np.sin(th))
    lat.upload("pi1", np.zeros(shape))
    lat.upload("pi2", np.zeros(shape))
    lat.set_table(t, tmin, dinv, odd=False)
    lat.step(10, timed=True, **kw)
    ms = lat.step(steps, timed=True, **kw)
    us = 1e3 * ms / steps
    t0 = time.perf_counter()
    for _ in range(5):
        strings, plaq = lat.winding_counts()
    w_ms = 1e3 * (time.perf_counter() - t0) / 5
    # past the roll (steps 10 .. 10 + steps include it): the state of most of a long run
    lat.step(3000 - 10 - steps, **kw)
    us_settled = 1e3 * lat.step(steps, timed=True, **kw) / steps
    out["complex256"] = {"us_per_step": us, "site_updates_per_s": n ** 3 / (us * 1e-6),
                         "hbm_gbps_algorithmic": 32.0 * n ** 3 / (us * 1e-6) / 1e9,
                         "bytes_per_site": 32.0, "us_per_step_settled": us_settled,
                         "hbm_gbps_algorithmic_settled": 32.0 * n ** 3 / (us_settled * 1e-6) / 1e9,
                         "winding_ms_wall": w_ms,
                         "winding_gbps_algorithmic": 11.0 * n ** 3 / (w_ms * 1e-3) / 1e9,
                         "strings": strings}
    lat.close()
    log(f"complex 256^3: {us:.1f} us/step (steps 10-{10 + steps}), {us_settled:.1f} us/step "
        f"after 3000 steps, winding {w_ms:.2f} ms")

    table, tmin, dinv, prm = real_workload(n)
    lat = Lattice(shape, np.float64)
    lat.upload("phi", 0.01 * np.random.default_rng(0).normal(size=shape))
    lat.upload("pi", np.zeros(shape))
    lat.set_table(table, tmin, dinv, odd=True)
    kw64 = dict(dt=prm["dt"], dx=prm["dx"], eta_eff=prm["eta_eff"], mu=prm["mu"],
                noise_scale=prm["noise_scale"])
    lat.step(5, timed=True, **kw64)
    ms = lat.step(50, timed=True, **kw64)
    us = 1e3 * ms / 50
    out["real256_f64"] = {"us_per_step": us, "site_updates_per_s": n ** 3 / (us * 1e-6),
                          "hbm_gbps_algorithmic": 32.0 * n ** 3 / (us * 1e-6) / 1e9,
                          "bytes_per_site": 32.0}
    lat.close()
    log(f"real 256^3 float64: {us:.1f} us/step")

    m = 512
    shape = (m, m, m)
    sur = S.Surrogate(S.SurrogateSpec(), seed=0)
    lat = Lattice(shape, np.float32, complex_field=True)
    for k in ("phi1", "phi2", "pi1", "pi2"):
        lat.upload(k, np.zeros(shape, np.float32) + (1.0 if k == "phi1" else 0.0))
    kw4 = dict(dt=dt, dx=dx, eta_eff=eta, mu=mu, noise_scale=kw["noise_scale"])
    dT = (6770.0 - 7350.0) / 1e5
    lat.step_mlp(2, sur, T_start=7350.0, dT_per_step=dT, timed=True, **kw4)
    nst = 10
    ms = lat.step_mlp(nst, sur, T_start=7350.0, dT_per_step=dT, timed=True, **kw4)
    us = 1e3 * ms / nst
    flop = 2 * 2 * (2 * 32 + 32 * 32 + 32) * m ** 3  # two RK2 stages, FMA = 2 flop
    lat.step_mlp(2, sur, T_start=7350.0, dT_per_step=dT, grid_points=16385, timed=True, **kw4)
    ms_g = lat.step_mlp(nst, sur, T_start=7350.0, dT_per_step=dT, grid_points=16385,
                        timed=True, **kw4)
    us_g = 1e3 * ms_g / nst
    out["surrogate512"] = {"us_per_step": us, "site_updates_per_s": m ** 3 / (us * 1e-6),
                           "mlp_tflops": flop / (us * 1e-6) / 1e12,
                           "flop_per_site_step": flop / m ** 3,
                           "hbm_gbps_algorithmic": 32.0 * m ** 3 / (us * 1e-6) / 1e9,
                           "grid_mode_us_per_step": us_g,
                           "grid_mode_site_updates_per_s": m ** 3 / (us_g * 1e-6),
                           "grid_mode_hbm_gbps_algorithmic": 32.0 * m ** 3 / (us_g * 1e-6) / 1e9}
    lat.close()
    log(f"complex 512^3 with the inline MLP: {us:.1f} us/step; per-step grid: {us_g:.1f} us/step")
    out.update(bench_other_integrators(steps))
    out["labels256"] = bench_labels()
    out["vprime_table"] = bench_vprime_table()
    return out


def bench_other_integrators(steps: int = 200) -> dict:
    """The reference's BAOAB and overdamped integrators at 256^3 (real and complex float32) on
    a resident lattice: one pass per BAOAB step (16 / 32 B per site), the overdamped step
    reads and writes phi only (8 / 16 B per site)."""
    from tip import potential as P
    from tip.lattice import Lattice

    out = {}
    n = 256
    shape = (n, n, n)
    table, tmin, dinv, prm = real_workload(n)
    rng = np.random.default_rng(5)
    phi0 = (0.01 * rng.normal(size=shape)).astype(np.float32)
    pc = P.ModelParams.complex()
    tc, tcmin, tcdinv = P.build_vprime_table(6770.0, 0.0, 1e5, 16384, pc)
    th = 2 * np.pi * rng.random(shape)
    r = 50.0 * np.abs(rng.normal(size=shape))
    for field in ("real", "complex"):
        for kind in ("baoab", "overdamped"):
            cplx = field == "complex"
            lat = Lattice(shape, np.float32, complex_field=cplx)
            if cplx:
                lat.upload("phi1", (r * np.cos(th)).astype(np.float32))
                lat.upload("phi2", (r * np.sin(th)).astype(np.float32))
                lat.upload("pi1", np.zeros(shape, np.float32))
                lat.upload("pi2", np.zeros(shape, np.float32))
                lat.set_table(tc, tcmin, tcdinv)
            else:
                lat.upload("phi", phi0)
                lat.upload("pi", np.zeros(shape, np.float32))
                lat.set_table(table, tmin, dinv, odd=True)
            kw = dict(dt=prm["dt"], dx=prm["dx"], mu=prm["mu"], noise_scale=30.0, seed=1)
            if kind == "baoab":
                run = lambda k: lat.step_baoab(k, gamma=prm["eta_eff"], timed=True, **kw)
            else:
                run = lambda k: lat.step_overdamped(k, eta_eff=prm["eta_eff"], timed=True, **kw)
            run(10)
            us = 1e3 * run(steps) / steps
            lat.close()
            b = (16.0 if kind == "baoab" else 8.0) * (2 if cplx else 1)
            passes = 1.0  # BAOAB: one fused pass per step (the last half-kick is deferred)
            out[f"{kind}_{field}256"] = {
                "us_per_step": us, "site_updates_per_s": n ** 3 / (us * 1e-6),
                "bytes_per_site_pass": b, "passes_per_step": passes,
                "hbm_gbps_algorithmic": b * passes * n ** 3 / (us * 1e-6) / 1e9}
            log(f"{kind} {field} 256^3: {us:.1f} us/step")
    return out


def bench_vprime_table() -> dict:
    """The V' table producer: tip_build_vprime_table on the device (8192 and 65536 points of
    the real model's table at 7000 GeV, the rebuild a Hubble-on run does every few steps)
    against the host producer (tip.potential.build_vprime_table, numpy float64)."""
    from tip import potential as P
    from tip import vprime_device as VD
    from tip.device import synchronize

    p = P.ModelParams.real()
    vt = P.thermal_vev(7000.0, p)
    out = {}
    for npts in (8192, 65536):
        tab, _, _ = VD.build_vprime_table(7000.0, 0.0, 1.5 * vt, npts, p)  # warm, spline upload
        synchronize()
        reps = 50
        t0 = time.perf_counter()
        for k in range(reps):
            VD.build_vprime_table(7000.0 - 0.01 * k, 0.0, 1.5 * vt, npts, p, out=tab)
        synchronize()
        dev_us = 1e6 * (time.perf_counter() - t0) / reps
        t0 = time.perf_counter()
        P.build_vprime_table(7000.0, 0.0, 1.5 * vt, npts, p)
        host_us = 1e6 * (time.perf_counter() - t0)
        tab.free()
        out[str(npts)] = {"device_us_wall": dev_us, "host_numpy_us": host_us}
        log(f"V' table {npts} points: device {dev_us:.1f} us (wall, incl. launch), host "
            f"{host_us:.0f} us")
    return out


def bench_labels() -> dict:
    """Component labelling of |phi| > esc at 256^3 (a smoothed field near the percolation
    threshold): the device labels and measures against scipy.ndimage.label plus the
    reference's numpy measures on the host cores."""
    from scipy import ndimage

    from tip import analysis as A
    from tip.device import DeviceArray

    n = 256
    rng = np.random.default_rng(9)
    f = ndimage.gaussian_filter(rng.normal(size=(n, n, n)), 1.5, mode="wrap")
    phi = (f / f.std() * 1e4).astype(np.float32)
    d = DeviceArray.from_numpy(phi)
    A.identify_bubbles(d, 5e3)  # warm
    t0 = time.perf_counter()
    reps = 5
    for _ in range(reps):
        comp = A.Components(d, 5e3, stats=False)
        comp.close()
    lab_ms = 1e3 * (time.perf_counter() - t0) / reps
    t0 = time.perf_counter()
    for _ in range(reps):
        bub = A.identify_bubbles(d, 5e3)
    all_ms = 1e3 * (time.perf_counter() - t0) / reps
    d.free()
    t0 = time.perf_counter()
    lab, nc = ndimage.label((phi > 5e3) | (phi < -5e3))
    sp_ms = 1e3 * (time.perf_counter() - t0)
    log(f"labels 256^3: {len(bub)} components, device labelling {lab_ms:.2f} ms, "
        f"with measures {all_ms:.2f} ms; scipy.ndimage.label {sp_ms:.1f} ms")
    return {"components": len(bub), "label_ms_wall": lab_ms, "label_and_measures_ms_wall": all_ms,
            "scipy_label_ms": sp_ms, "scipy_components": int(nc)}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--workload", default="real256", choices=["real256", "real256_f64"])
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-ensemble", action="store_true")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the complex, float64 and surrogate throughput lines")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-config-runs", action="store_true",
                    help="skip the end-to-end 256^3 runs at their own lengths (~10 s)")
    ap.add_argument("--ensemble-modes", default="graph,batched,sequential",
                    help="comma-separated subset of graph, batched, sequential")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="after the run, 5 device-to-device copies of 128 MiB (a known-size "
                         "reference for the HBM byte counters of a profiler)")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (--gpus {a.gpus} requested); the "
              "multi-GPU sharding is not implemented.", file=sys.stderr)
        return 2

    import tip
    from tip.lattice import Lattice

    if tip.device_count() < 1:
        print("bench.py: no GPU visible", file=sys.stderr)
        return 1
    n = 256
    shape = (n, n, n)
    dtype = np.float64 if a.workload == "real256_f64" else np.float32
    table, tmin, dinv, prm = real_workload(n)
    log(f"workload {a.workload}: {shape} {np.dtype(dtype).name}, table {len(table)} pts, "
        f"ns={prm['noise_scale']:.4g}")
    lat = Lattice(shape, dtype)
    rng = np.random.default_rng(0)
    lat.upload("phi", 0.01 * rng.normal(size=shape))
    lat.upload("pi", np.zeros(shape))
    lat.set_table(table, tmin, dinv, odd=True)
    kw = dict(dt=prm["dt"], dx=prm["dx"], eta_eff=prm["eta_eff"], mu=prm["mu"],
              inv_a2=prm["inv_a2"], noise_scale=prm["noise_scale"])
    lat.step(a.warmup, timed=True, **kw)
    log(f"warmup {a.warmup} steps done")
    t0 = time.perf_counter()
    ev_ms = lat.step(a.steps, timed=True, **kw)  # synchronous (event sync) on return
    wall = time.perf_counter() - t0
    log(f"timed {a.steps} steps: wall {wall * 1e3:.2f} ms, events {ev_ms:.2f} ms")
    frac_true = lat.vacuum_fraction(1e4)
    lo, hi = lat.minmax()
    if not (math.isfinite(lo) and math.isfinite(hi)):
        log("non-finite field after the run")
        return 1

    sites = float(n ** 3)
    bps = 16.0 if dtype == np.float32 else 32.0
    value = sites * a.steps / wall
    kern_s = ev_ms * 1e-3 / a.steps
    achieved = sites * bps / kern_s / 1e9
    prof_name = "rk2_fused_f32_256cube.md" if dtype == np.float32 else "rk2_fused_f64_256cube.md"
    prof = profile_record(prof_name)
    traffic = prof.get("traffic_bytes_per_launch")
    out = {
        "metric": "site-updates/s + HBM GB/s at 256^3 FP32; batched ensemble vs plain loop",
        "value": value,
        "unit": "site-updates/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": 1e3 * wall / a.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": value / REF_CPU_SITE_UPDATES,
        "dtype": "f32" if dtype == np.float32 else "f64",
        "data": "synthetic",
        "config": {"workload": f"real scalar 256^3 periodic, {np.dtype(dtype).name}, fused RK2 "
                               f"+ inline hash noise, 8192-pt V' table at T0=7000",
                   "lattice": [n, n, n], "dt": prm["dt"], "dx": prm["dx"],
                   "eta": prm["eta_eff"], "table_points": len(table)},
        "hbm_gbps_algorithmic": sites * bps * a.steps / wall / 1e9,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBPS, "traffic": traffic,
                     "kernel_us": kern_s * 1e6, "bytes_per_site": bps,
                     # traffic is not measured in this run: it is the committed rocprofv3
                     # FETCH_SIZE + WRITE_SIZE record of this workload, whose own kernel
                     # average is traffic_kernel_us
                     "traffic_source": f"profiles/{prof_name}" if traffic else None,
                     "traffic_kernel_us": prof.get("kernel_average_us")},
        "true_vacuum_fraction": frac_true,
    }
    settled = None
    if not a.no_extras and dtype == np.float32:
        # the same lattice past its roll to the true vacuum (the state most of a 10^4-step run
        # spends its time in): the step's table gathers then spread over many cache lines
        done = a.warmup + a.steps
        if done < 3000:
            lat.step(3000 - done, **kw)
        ms = lat.step(200, timed=True, **kw)
        us = 1e3 * ms / 200
        settled = {"after_steps": max(done, 3000), "us_per_step": us,
                   "site_updates_per_s": sites / (us * 1e-6),
                   "hbm_gbps_algorithmic": sites * bps / (us * 1e-6) / 1e9,
                   "frac": sites * bps / (us * 1e-6) / 1e9 / HBM_PEAK_GBPS,
                   "true_vacuum_fraction": lat.vacuum_fraction(1e4)}
        log(f"real 256^3 after the roll (steps {settled['after_steps']}+200): {us:.1f} us/step")
    lat.close()
    if a.copy_calibration:
        from tip.device import DeviceArray, synchronize
        from tip import _lib
        nbytes = 128 << 20
        src = DeviceArray((nbytes // 4,), np.float32)
        dst = DeviceArray((nbytes // 4,), np.float32)
        src.fill_zero()
        for _ in range(5):
            _lib.call("tip_memcpy_d2d", dst.ptr, src.ptr, nbytes, None)
        synchronize()
        src.free()
        dst.free()
    if not a.no_ensemble:
        log("batched ensemble (64 realisations, 46.4M sites) ...")
        out["ensemble"] = bench_ensemble(modes=[m for m in a.ensemble_modes.split(",") if m])
    if not a.no_extras:
        out["extras"] = bench_extras()
        if not a.no_config_runs:
            ens = out.get("ensemble")
            settings = ((ens["autotune_best"], ens["group_bytes"], ens["groups_at_once"])
                        if ens else None)
            out["extras"]["config_runs"] = bench_config_runs(settings)
        if settled:
            out["extras"]["real256_settled"] = settled
    if not a.no_cpu:
        log("cpu baseline ...")
        out["cpu_baseline"] = cpu_baseline(shape, table, tmin, dinv, prm, a.cpu_seconds)
        out["cpu_baseline"]["host"] = host_description()
        t64, tmin64, dinv64, prm64 = real_workload(64)
        leg = cpu_baseline((64, 64, 64), t64, tmin64, dinv64, prm64, a.cpu_seconds / 4)
        out["cpu_baseline"]["legs"] = {"real64_f32": leg}
        log(f"cpu 64^3: {leg['ms_per_step']:.2f} ms/step")
        if not a.no_ensemble:
            leg = cpu_baseline_ensemble(a.cpu_seconds / 2)
            out["cpu_baseline"]["legs"]["nucleation_ensemble_sequential"] = leg
            log(f"cpu ensemble: {leg['ms_per_step']:.1f} ms/step")
        if "extras" in out:
            out["extras"]["cpu_baseline_integrators"] = cpu_baseline_integrators(
                table, tmin, dinv, prm)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
