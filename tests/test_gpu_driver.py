"""The lattice driver on the GPU (-m gpu): its main loop replays exactly on the oracle, and the
surrogate path's multi-step calls equal one call per step.

Bars:
  * real float64 with Hubble cooling, 30 steps at 32^3, tables rebuilt by the driver's rule on
    the device (no host table build): fields within 1e-12 relative of the oracle fed the host
    producer's tables for the same requests, schedule and noise seeds; true-vacuum series exact
  * complex float32 with the surrogate under linear cooling: one call per report interval
    (noise amplitude scaled by sqrt(T_n / T_start) inside) within 1e-5 relative of one call per
    step (noise amplitude from the formula at every step)
"""
import math

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def test_driver_real_f64_replays_on_oracle(monkeypatch):
    from tip import driver as D
    from tip import potential as P
    cfg = D.RunConfig(field="real", Nx=32, Ny=32, Nz=32, Nt=30, T0=7000.0, precision="f64",
                      no_hubble=False, report=10, seed=3, quiet=True, table_size=4096)
    run = D.Run(cfg)
    builds = []  # (step index the table is first used at, T, phi_min, phi_max, npts)
    orig = run.lat.build_table

    def rec(T, lo, hi, n, model, spline=None, odd=False):
        builds.append((run.lat.step_count, T, lo, hi, n))
        return orig(T, lo, hi, n, model, spline, odd=odd)

    def host_build(*a, **k):
        raise AssertionError("the driver built a V' table on the host")

    run.lat.build_table = rec
    monkeypatch.setattr(P, "build_vprime_table", host_build)
    try:
        series = run.run()
        phi_g, pi_g = run.lat.download("phi"), run.lat.download("pi")
    finally:
        run.close()
        monkeypatch.undo()
    assert len(builds) >= 1 and series["step"] == [10, 20, 30]
    # the oracle is fed the host producer's tables of the same (T, range) requests
    tables = [(n0, *P.build_vprime_table(T, lo, hi, n, run.model, run.spline))
              for n0, T, lo, hi, n in builds]

    f0 = D.initial_fields(cfg)
    phi, pi = f0["phi"].copy(), f0["pi"].copy()
    pt, qt = np.empty_like(phi), np.empty_like(pi)
    hub = P.HubbleState(cfg.T0, D.MU, cfg.eta_phys / D.MU)
    dt, dx = D.MU * cfg.dt_phys, D.MU * cfg.dx_phys
    fr = []
    k = 0
    for n in range(cfg.Nt):
        while k + 1 < len(tables) and tables[k + 1][0] <= n:
            k += 1
        _, t, tmin, dinv = tables[k]
        s = hub.step(dt)
        ns = math.sqrt(2.0 * s["eta_eff"] * s["T"] * dt * s["inv_a3"] / (D.MU ** 2 * cfg.dx_phys ** 3))
        noise = O.noise_field(phi.shape, O.noise_seed(n, cfg.seed), ns)
        O.rk2_step_table(phi, pi, dt, dx, s["eta_eff"], D.MU, noise, pt, qt, s["inv_a2"], t,
                         tmin, dinv, t.size)
        if (n + 1) % cfg.report == 0:
            fr.append(O.vacuum_count(phi, cfg.esc) / phi.size)
    assert np.linalg.norm(phi_g - phi) <= 1e-12 * np.linalg.norm(phi)
    assert np.linalg.norm(pi_g - pi) <= 1e-12 * np.linalg.norm(pi)
    assert series["f_true"] == fr


def test_surrogate_chunked_equals_stepwise():
    from tip import driver as D
    from tip import surrogate as S
    sur = S.Surrogate(S.SurrogateSpec(), params=S.init_params(11))
    out = []
    for chunked in (True, False):
        cfg = D.RunConfig(field="complex", Nx=32, Ny=32, Nz=64, Nt=40, T0=7350.0,
                          cool_to=6770.0, report=20, seed=5, quiet=True, surrogate=sur,
                          init_rho=50.0)
        run = D.Run(cfg)
        if not chunked:
            run._chunk = lambda n: 1
        try:
            run.run()
            out.append([run.lat.download(k).astype(np.float64)
                        for k in ("phi1", "phi2", "pi1", "pi2")])
        finally:
            run.close()
    for a, b in zip(*out):
        assert np.linalg.norm(a - b) <= 1e-5 * np.linalg.norm(b)


@pytest.mark.parametrize("shape", [(32, 32, 64), (5, 7, 300)])
def test_save_point_on_device(shape):
    """The complex save point computed on the device (latticeSimComplex_numba.py:4625-4646):
    rho and its range bit-exact against numpy's float64 formula, theta within one float32 ulp
    of numpy's arctan2, the winding buffer equal to the oracle's rounded site sums."""
    from tip.lattice import Lattice
    rng = np.random.default_rng(shape[2])
    p1 = (rng.normal(size=shape) * 300.0).astype(np.float32)
    p2 = (rng.normal(size=shape) * 300.0).astype(np.float32)
    p1.ravel()[:4] = [0.0, -0.0, 0.0, 1e-30]
    p2.ravel()[:4] = [0.0, 0.0, -0.0, -1e-30]
    lat = Lattice(shape, np.float32, complex_field=True)
    try:
        lat.upload("phi1", p1)
        lat.upload("phi2", p2)
        sp = lat.save_point()
    finally:
        lat.close()
    rho = np.sqrt(p1.astype(np.float64) ** 2 + p2.astype(np.float64) ** 2)
    theta = np.arctan2(p2.astype(np.float64), p1.astype(np.float64)).astype(np.float32)
    assert np.array_equal(sp["rho"], rho.astype(np.float32))
    assert sp["rho_min"] == float(rho.min()) and sp["rho_max"] == float(rho.max())
    ulp = np.abs(sp["theta"].view(np.int32).astype(np.int64) - theta.view(np.int32))
    assert int(ulp.max()) <= 1
    assert np.array_equal(np.signbit(sp["theta"]), np.signbit(theta))
    w_ref, _ = O.winding(p1, p2)
    assert np.array_equal(sp["winding"], np.rint(w_ref).astype(np.float32))
    assert sp["n_strings"] == O.string_count(w_ref)


def test_driver_snapshot_keys(tmp_path):
    """A complex run with snapshots writes the reference's npz keys; the derived arrays agree
    with the saved fields."""
    from tip import driver as D
    cfg = D.RunConfig(field="complex", Nx=16, Ny=16, Nz=64, Nt=20, T0=7350.0, report=10,
                      steps=10, seed=2, quiet=True, outdir=str(tmp_path), init_rho=50.0)
    run = D.Run(cfg)
    try:
        run.run()
    finally:
        run.close()
    files = sorted(tmp_path.glob("state_step_*.npz"))
    assert len(files) >= 2
    with np.load(files[-1]) as z:
        keys = set(z.files)
        assert {"phi1", "phi2", "rho", "theta", "winding", "step", "time", "temperature",
                "rho_min", "rho_max", "zn_active"} <= keys
        p1, p2 = z["phi1"], z["phi2"]
        rho = np.sqrt(p1.astype(np.float64) ** 2 + p2.astype(np.float64) ** 2)
        assert np.array_equal(z["rho"], rho.astype(np.float32))
        assert float(z["rho_max"]) == float(rho.max())
        w_ref, _ = O.winding(p1, p2)
        assert np.array_equal(z["winding"], np.rint(w_ref).astype(np.float32))


@pytest.mark.parametrize("integrator", ["baoab", "overdamped"])
def test_driver_real_f64_other_integrators_replay(integrator):
    """The real float64 driver with Hubble expansion (inv_a2, eta_eff and T change every step)
    under BAOAB and overdamped: fields within 1e-12 of the oracle fed the same tables and
    schedule (BAOAB's deferred half-kick keeps each step's own inv_a2)."""
    from tip import driver as D
    from tip import potential as P
    cfg = D.RunConfig(field="real", Nx=24, Ny=20, Nz=32, Nt=24, T0=7000.0, precision="f64",
                      no_hubble=False, report=8, seed=4, quiet=True, table_size=4096,
                      integrator=integrator)
    run = D.Run(cfg)
    builds = []
    orig = run.lat.build_table

    def rec(T, lo, hi, n, model, spline=None, odd=False):
        builds.append((run.lat.step_count, T, lo, hi, n))
        return orig(T, lo, hi, n, model, spline, odd=odd)

    run.lat.build_table = rec
    try:
        run.run()
        phi_g, pi_g = run.lat.download("phi"), run.lat.download("pi")
    finally:
        run.close()
    tables = [(n0, *P.build_vprime_table(T, lo, hi, n, run.model, run.spline))
              for n0, T, lo, hi, n in builds]
    f0 = D.initial_fields(cfg)
    phi, pi = f0["phi"].copy(), f0["pi"].copy()
    hub = P.HubbleState(cfg.T0, D.MU, cfg.eta_phys / D.MU)
    dt, dx = D.MU * cfg.dt_phys, D.MU * cfg.dx_phys
    k = 0
    for n in range(cfg.Nt):
        while k + 1 < len(tables) and tables[k + 1][0] <= n:
            k += 1
        _, t, tmin, dinv = tables[k]
        s = hub.step(dt)
        dx3 = D.MU ** 2 * cfg.dx_phys ** 3
        if integrator == "overdamped":
            ns = math.sqrt(2.0 * s["T"] * dt * s["inv_a3"] / (s["eta_eff"] * dx3))
            phi, _ = O.langevin(O.OVERDAMPED, phi, None, dt, dx, s["eta_eff"], D.MU,
                                s["inv_a2"], t, tmin, dinv, noise_seed=O.noise_seed(n, cfg.seed),
                                noise_scale=ns)
        else:
            ns = math.sqrt(s["T"] * s["inv_a3"] / dx3)
            phi, pi = O.baoab_step(phi, pi, dt, dx, s["eta_eff"], D.MU, s["inv_a2"], t, tmin,
                                   dinv, noise_seed=O.noise_seed(n, cfg.seed), noise_scale=ns)
    assert np.linalg.norm(phi_g - phi) <= 1e-12 * np.linalg.norm(phi)
    if integrator == "baoab":
        assert np.linalg.norm(pi_g - pi) <= 1e-12 * np.linalg.norm(pi)
    else:
        assert np.array_equal(pi_g, f0["pi"])


@pytest.mark.parametrize("field,integrator", [("real", "rk2"), ("complex", "rk2"),
                                              ("real", "baoab")])
def test_range_every_chunks_equal_stepwise(field, integrator, monkeypatch):
    """Constant parameters with the table range checked every 25 steps (steps in between sent
    to the GPU in one call): the same fields, bit for bit, and the same table builds as the
    reference's check at every step, on a run whose field stays inside the first table (a
    table margin of at least 10^6 GeV keeps it there)."""
    from tip import driver as D
    monkeypatch.setattr(D, "TABLE_MARGIN_MIN", 1.0e6)
    outs = []
    for every in (1, 25):
        cfg = D.RunConfig(field=field, Nx=24, Ny=20, Nz=64, Nt=100, T0=7000.0 if field == "real"
                          else 6770.0, integrator=integrator, report=50, seed=5, quiet=True,
                          table_size=2048, range_every=every)
        run = D.Run(cfg)
        try:
            series = run.run()
            names = ("phi1", "phi2", "pi1", "pi2") if field == "complex" else ("phi", "pi")
            outs.append(([run.lat.download(k) for k in names], run.rebuilds, series))
        finally:
            run.close()
    (a, ra, sa), (b, rb, sb) = outs
    assert ra == rb == 1
    assert all(np.array_equal(x, y) for x, y in zip(a, b))
    assert sa["f_true"] == sb["f_true"] and sa["step"] == sb["step"] == [50, 100]
