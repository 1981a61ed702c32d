"""Overdamped Euler-Maruyama and BAOAB on the GPU (-m gpu) against the CPU oracle.

Reference: overdamped_euler_step_table / baoab_step_table,
simulation/latticeSimeRescale_numba.py:1974-2192 and simulation/latticeSimComplex_numba.py:
2223-2517, in the race-free form the oracle restates (oracle/tip_oracle.c orc_langevin_*).

Bars:
  * float64 drop-ins (real; complex with Z_N on): fields within 1e-12 relative of the oracle
    after 6 steps with hash noise (the only differences are libm's log/cos/sin/atan2)
  * float32 drop-ins: within 2e-5 relative of the float64 oracle after 6 steps
  * lattice BAOAB over k steps (open, fused x (k-1), deferred close): float64 bit-identical
    to k drop-in steps, also as k single-step calls across table and inv_a2 changes; float32
    within 1e-6 relative; the overdamped lattice path equals the drop-in loop and leaves pi
    untouched
  * the driver's complex overdamped and BAOAB loops replay on the oracle fed the same tables
    (float32 bar 2e-5)
"""
import math

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

SHAPES = [(12, 10, 70), (5, 7, 33)]


def rel(a, b):
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))


def real_setup(shape, seed):
    from tip import potential as P
    rng = np.random.default_rng(seed)
    p = P.ModelParams.real()
    vt = P.thermal_vev(7000.0, p)
    t, tmin, dinv = P.build_vprime_table(7000.0, 0.0, 1.5 * vt, 4096, p)
    phi = rng.normal(size=shape) * 0.3 * vt
    pi = rng.normal(size=shape) * 50.0
    # dt, dx, eta / gamma, mu, inv_a2, noise_scale
    prm = dict(dt=0.25, dx=2.0, eta=7.0, mu=1000.0, inv_a2=0.93, ns=40.0)
    return phi, pi, (t, tmin, dinv), prm


def complex_setup(shape, seed):
    from tip import potential as P
    rng = np.random.default_rng(seed)
    t, tmin, dinv = P.build_vprime_table(6770.0, 0.0, 1e5, 4096, P.ModelParams.complex())
    f = [rng.normal(size=shape) * 2e3 for _ in range(2)]
    q = [rng.normal(size=shape) * 1e2 for _ in range(2)]
    prm = dict(dt=0.25, dx=1.0, eta=1.3, mu=1000.0, inv_a2=1.0, ns=300.0)
    return f, q, (t, tmin, dinv), prm


def oracle_real(op, phi, pi, tab, prm, nsteps, odd=True):
    for n in range(nsteps):
        kw = dict(noise_seed=O.noise_seed(n), noise_scale=prm["ns"])
        if op == "overdamped":
            phi, _ = O.langevin(O.OVERDAMPED, phi, None, prm["dt"], prm["dx"], prm["eta"],
                                prm["mu"], prm["inv_a2"], *tab, odd=odd, **kw)
        else:
            phi, pi = O.baoab_step(phi, pi, prm["dt"], prm["dx"], prm["eta"], prm["mu"],
                                   prm["inv_a2"], *tab, odd=odd, **kw)
    return phi, pi


def oracle_complex(op, f, q, tab, prm, zn, nsteps):
    a, b = f
    u, v = q
    for n in range(nsteps):
        kw = dict(zn=zn, noise_seed=O.noise_seed(n), noise_scale=prm["ns"])
        args = (prm["dt"], prm["dx"], prm["eta"], prm["mu"], prm["inv_a2"], *tab)
        if op == "overdamped":
            a, b, _, _ = O.langevin_complex(O.OVERDAMPED, a, b, None, None, *args, **kw)
        else:
            a, b, u, v = O.langevin_complex(O.BAOAB_OPEN, a, b, u, v, *args, **kw)
            _, _, u, v = O.langevin_complex(O.BAOAB_CLOSE, a, b, u, v, *args[:5], *tab, zn=zn)
    return (a, b), (u, v)


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("op", ["overdamped", "baoab"])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_real_drop_in_matches_oracle(shape, op, dtype):
    from tip import kernels as K
    phi0, pi0, tab, prm = real_setup(shape, 3)
    phi0, pi0 = phi0.astype(dtype), pi0.astype(dtype)
    phi, pi = phi0.copy(), pi0.copy()
    for n in range(6):
        if op == "overdamped":
            K.overdamped_euler_step_table(phi, prm["dt"], prm["dx"], prm["eta"], prm["mu"],
                                          prm["inv_a2"], tab[0], tab[1], tab[2], len(tab[0]),
                                          prm["ns"], O.noise_seed(n), odd=True)
        else:
            K.baoab_step_table(phi, pi, prm["dt"], prm["dx"], prm["eta"], prm["mu"],
                               prm["inv_a2"], tab[0], tab[1], tab[2], len(tab[0]), prm["ns"],
                               O.noise_seed(n), odd=True)
    rp, rq = oracle_real(op, phi0.astype(np.float64), pi0.astype(np.float64), tab, prm, 6)
    bar = 1e-12 if dtype == np.float64 else 2e-5
    assert rel(phi.astype(np.float64), rp) < bar
    if op == "baoab":
        assert rel(pi.astype(np.float64), rq) < bar
    else:
        assert np.array_equal(pi, pi0)


@pytest.mark.parametrize("zn", [(3.0, 1e8, 1.0), (3.0, 1e8, 0.0)])
@pytest.mark.parametrize("op", ["overdamped", "baoab"])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_complex_drop_in_matches_oracle(op, dtype, zn):
    from tip import kernels_complex as KC
    shape = (10, 9, 70)
    f, q, tab, prm = complex_setup(shape, 5)
    a, b = (x.astype(dtype) for x in f)
    u, v = (x.astype(dtype) for x in q)
    f0 = [a.astype(np.float64), b.astype(np.float64)]
    q0 = [u.astype(np.float64), v.astype(np.float64)]
    for n in range(6):
        if op == "overdamped":
            KC.overdamped_euler_step_table(a, b, prm["dt"], prm["dx"], prm["eta"], prm["mu"],
                                           prm["inv_a2"], tab[0], tab[1], tab[2], len(tab[0]),
                                           prm["ns"], O.noise_seed(n), zn=zn)
        else:
            KC.baoab_step_table(a, b, u, v, prm["dt"], prm["dx"], prm["eta"], prm["mu"],
                                prm["inv_a2"], tab[0], tab[1], tab[2], len(tab[0]), prm["ns"],
                                O.noise_seed(n), zn=zn)
    (ra, rb), (ru, rv) = oracle_complex(op, f0, q0, tab, prm, zn, 6)
    bar = 1e-12 if dtype == np.float64 else 2e-5
    assert rel(a.astype(np.float64), ra) < bar and rel(b.astype(np.float64), rb) < bar
    if op == "baoab":
        assert rel(u.astype(np.float64), ru) < bar and rel(v.astype(np.float64), rv) < bar


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_lattice_baoab_fused_equals_steps(dtype):
    from tip import kernels as K
    from tip.lattice import Lattice
    shape = (16, 12, 70)
    phi0, pi0, tab, prm = real_setup(shape, 7)
    phi0, pi0 = phi0.astype(dtype), pi0.astype(dtype)
    k = 9
    lat = Lattice(shape, dtype)
    try:
        lat.upload("phi", phi0)
        lat.upload("pi", pi0)
        lat.set_table(*tab, odd=True)
        lat.step_baoab(k, dt=prm["dt"], dx=prm["dx"], gamma=prm["eta"], mu=prm["mu"],
                       inv_a2=prm["inv_a2"], noise_scale=prm["ns"], seed=0, step0=0)
        lp, lq = lat.download("phi"), lat.download("pi")
    finally:
        lat.close()
    phi, pi = phi0.copy(), pi0.copy()
    for n in range(k):
        K.baoab_step_table(phi, pi, prm["dt"], prm["dx"], prm["eta"], prm["mu"], prm["inv_a2"],
                           tab[0], tab[1], tab[2], len(tab[0]), prm["ns"], O.noise_seed(n),
                           odd=True)
    if dtype == np.float64:
        assert np.array_equal(lp, phi) and np.array_equal(lq, pi)
    else:
        assert rel(lp.astype(np.float64), phi.astype(np.float64)) < 1e-6
        assert rel(lq.astype(np.float64), pi.astype(np.float64)) < 1e-6


def test_lattice_baoab_deferred_kick():
    """Single-step BAOAB calls fuse the deferred half-kick into the next call: k calls of one
    step equal one call of k steps and the drop-in loop (float64, bitwise), also across a
    table change (the pending kick runs with the old table first) and an inv_a2 change
    (the fused pass takes both); reading phi does not complete the kick, reading pi does."""
    from tip import kernels as K
    from tip.lattice import Lattice
    shape = (12, 10, 70)
    phi0, pi0, tab, prm = real_setup(shape, 11)
    tab2 = (tab[0] * 1.01, tab[1], tab[2])
    kw = dict(dt=prm["dt"], dx=prm["dx"], gamma=prm["eta"], mu=prm["mu"],
              noise_scale=prm["ns"], seed=0)
    inv = [0.93, 0.93, 0.95, 0.97, 0.97, 0.97]
    lat = Lattice(shape, np.float64)
    try:
        lat.upload("phi", phi0)
        lat.upload("pi", pi0)
        lat.set_table(*tab, odd=True)
        for n in range(6):
            if n == 4:
                lat.set_table(*tab2, odd=True)
            lat.step_baoab(1, inv_a2=inv[n], step0=n, **kw)
            if n == 2:
                mid_phi = lat.download("phi")  # no completion
        lp, lq = lat.download("phi"), lat.download("pi")
    finally:
        lat.close()
    phi, pi = phi0.copy(), pi0.copy()
    for n in range(6):
        t = tab2 if n >= 4 else tab
        K.baoab_step_table(phi, pi, prm["dt"], prm["dx"], prm["eta"], prm["mu"], inv[n],
                           t[0], t[1], t[2], len(t[0]), prm["ns"], O.noise_seed(n), odd=True)
        if n == 2:
            assert np.array_equal(mid_phi, phi)
    assert np.array_equal(lp, phi) and np.array_equal(lq, pi)


def test_lattice_complex_integrators_match_drop_in():
    from tip import kernels_complex as KC
    from tip.lattice import Lattice
    shape = (8, 12, 96)
    f, q, tab, prm = complex_setup(shape, 9)
    zn = (3.0, 1e8, 1.0)
    init = dict(phi1=f[0], phi2=f[1], pi1=q[0], pi2=q[1])
    init = {k: v.astype(np.float32) for k, v in init.items()}
    for op in ("overdamped", "baoab"):
        lat = Lattice(shape, np.float32, complex_field=True)
        try:
            for k_, v_ in init.items():
                lat.upload(k_, v_)
            lat.set_table(*tab)
            lat.set_zn(*zn)
            kw = dict(dt=prm["dt"], dx=prm["dx"], mu=prm["mu"], inv_a2=prm["inv_a2"],
                      noise_scale=prm["ns"], seed=0, step0=0)
            if op == "baoab":
                lat.step_baoab(5, gamma=prm["eta"], **kw)
            else:
                lat.step_overdamped(5, eta_eff=prm["eta"], **kw)
            got = {k_: lat.download(k_) for k_ in init}
        finally:
            lat.close()
        a, b, u, v = (init[k_].copy() for k_ in ("phi1", "phi2", "pi1", "pi2"))
        for n in range(5):
            if op == "baoab":
                KC.baoab_step_table(a, b, u, v, prm["dt"], prm["dx"], prm["eta"], prm["mu"],
                                    prm["inv_a2"], *tab, len(tab[0]), prm["ns"],
                                    O.noise_seed(n), zn=zn)
            else:
                KC.overdamped_euler_step_table(a, b, prm["dt"], prm["dx"], prm["eta"],
                                               prm["mu"], prm["inv_a2"], *tab, len(tab[0]),
                                               prm["ns"], O.noise_seed(n), zn=zn)
        for k_, ref in zip(("phi1", "phi2", "pi1", "pi2"), (a, b, u, v)):
            assert rel(got[k_].astype(np.float64), ref.astype(np.float64)) < 1e-6, (op, k_)
        if op == "overdamped":
            assert np.array_equal(got["pi1"], init["pi1"])


@pytest.mark.parametrize("integrator", ["overdamped", "baoab"])
def test_driver_complex_integrators_replay_on_oracle(integrator):
    from tip import driver as D
    cfg = D.RunConfig(field="complex", Nx=16, Ny=16, Nz=64, Nt=12, T0=7350.0, cool_to=7000.0,
                      report=6, seed=0, quiet=True, init_rho=300.0, integrator=integrator,
                      table_size=4096, zn_order=3, zn_strength=1e8)
    from tip import potential as P
    run = D.Run(cfg)
    builds = []  # the driver builds its tables on the device; replay them from the host's
    orig = run.lat.build_table

    def rec(T, lo, hi, n, model, spline=None, odd=False):
        builds.append((run.lat.step_count, T, lo, hi, n))
        return orig(T, lo, hi, n, model, spline, odd=odd)

    run.lat.build_table = rec
    try:
        run.run()
        got = [run.lat.download(k).astype(np.float64) for k in ("phi1", "phi2", "pi1", "pi2")]
    finally:
        run.close()
    tables = [(n0, *P.build_vprime_table(T, lo, hi, n, run.model, run.spline))
              for n0, T, lo, hi, n in builds]
    f0 = D.initial_fields(cfg)
    a, b = f0["phi1"].astype(np.float64), f0["phi2"].astype(np.float64)
    u, v = f0["pi1"].astype(np.float64), f0["pi2"].astype(np.float64)
    dt, dx, eta = D.MU * cfg.dt_phys, D.MU * cfg.dx_phys, cfg.eta_phys / D.MU
    k = 0
    for n in range(cfg.Nt):
        while k + 1 < len(tables) and tables[k + 1][0] <= n:
            k += 1
        _, t, tmin, dinv = tables[k]
        T = cfg.T0 + (cfg.cool_to - cfg.T0) * n / cfg.Nt
        zn = (3.0, 1e8, 1.0)
        if integrator == "overdamped":
            ns = math.sqrt(2.0 * T * dt / (eta * D.MU ** 2 * cfg.dx_phys ** 3))
            a, b, _, _ = O.langevin_complex(O.OVERDAMPED, a, b, None, None, dt, dx, eta, D.MU,
                                            1.0, t, tmin, dinv, zn=zn,
                                            noise_seed=O.noise_seed(n), noise_scale=ns)
        else:
            ns = math.sqrt(T / (D.MU ** 2 * cfg.dx_phys ** 3))
            a, b, u, v = O.langevin_complex(O.BAOAB_OPEN, a, b, u, v, dt, dx, eta, D.MU, 1.0, t,
                                            tmin, dinv, zn=zn, noise_seed=O.noise_seed(n),
                                            noise_scale=ns)
            _, _, u, v = O.langevin_complex(O.BAOAB_CLOSE, a, b, u, v, dt, dx, eta, D.MU, 1.0,
                                            t, tmin, dinv, zn=zn)
    for g, r in zip(got, (a, b, u, v)):
        assert rel(g, r) < 2e-5


def test_full_size_noise_statistics():
    """256^3 float32 with the force switched off (zero table, dx so large the Laplacian
    vanishes): the BAOAB momentum is an Ornstein-Uhlenbeck process, var(pi) after n steps =
    ns^2 (1 - c1^(2n)), c1 = exp(-gamma dt); the overdamped field is a random walk, var(phi)
    = n ns^2, with independent draws per step (lag-one correlation of the increments ~ 0).
    Bars: 1 % on the variances (the sampling error at 2^24 sites is 0.03 %)."""
    from tip.lattice import Lattice
    shape = (256, 256, 256)
    zero = np.zeros(shape, np.float32)
    tab = (np.zeros(9), 0.0, 1.0)
    ns, gamma, dt = 0.7, 0.4, 0.25
    lat = Lattice(shape, np.float32)
    try:
        lat.upload("phi", zero)
        lat.upload("pi", zero)
        lat.set_table(*tab)
        n = 12
        lat.step_baoab(n, dt=dt, dx=1e15, gamma=gamma, mu=1.0, noise_scale=ns, seed=3)
        pi = lat.download("pi").astype(np.float64)
        want = ns ** 2 * (1.0 - math.exp(-2.0 * gamma * dt * n))
        assert abs(pi.var() / want - 1.0) < 0.01 and abs(pi.mean()) < 0.01 * ns
        lat.upload("phi", zero)
        lat.step_overdamped(1, dt=dt, dx=1e15, eta_eff=1.0, mu=1.0, noise_scale=ns, seed=5,
                            step0=0)
        d1 = lat.download("phi").astype(np.float64)
        lat.step_overdamped(7, dt=dt, dx=1e15, eta_eff=1.0, mu=1.0, noise_scale=ns, seed=5,
                            step0=1)
        d8 = lat.download("phi").astype(np.float64)
        assert abs(d1.var() / ns ** 2 - 1.0) < 0.01 and abs(d8.var() / (8 * ns ** 2) - 1.0) < 0.01
        inc = d8 - d1
        assert abs(np.corrcoef(d1.ravel()[:1 << 22], inc.ravel()[:1 << 22])[0, 1]) < 3e-3
    finally:
        lat.close()


@pytest.mark.parametrize("complex_field", [False, True])
def test_lattice_integrator_geometries(complex_field):
    """The row tiles (1, 2, 4 y rows per thread, several plane counts) and the flat tile give
    the same BAOAB and overdamped fields to float32 rounding (the two forms differ only in the
    table index arithmetic at segment boundaries)."""
    from tip.lattice import Lattice
    shape = (20, 16, 96)
    if complex_field:
        f, q, tab, prm = complex_setup(shape, 13)
        init = dict(phi1=f[0], phi2=f[1], pi1=q[0], pi2=q[1])
    else:
        phi0, pi0, tab, prm = real_setup(shape, 13)
        init = dict(phi=phi0, pi=pi0)
    init = {k: v.astype(np.float32) for k, v in init.items()}
    kw = dict(dt=prm["dt"], dx=prm["dx"], mu=prm["mu"], inv_a2=prm["inv_a2"],
              noise_scale=prm["ns"], seed=2, step0=0)
    out = {}
    for geo in [(0, 0, 1, 2), (1, 0, 1, 1), (2, 3, 0, 1), (4, 7, 1, 1), (2, 40, 1, 1)]:
        for op in ("baoab", "overdamped"):
            lat = Lattice(shape, np.float32, complex_field=complex_field)
            try:
                for k, v in init.items():
                    lat.upload(k, v)
                lat.set_table(*tab, odd=not complex_field)
                if complex_field:
                    lat.set_zn(3.0, 1e8, 1.0)
                lat.set_geometry(*geo)
                if op == "baoab":
                    lat.step_baoab(5, gamma=prm["eta"], **kw)
                else:
                    lat.step_overdamped(5, eta_eff=prm["eta"], **kw)
                out[geo, op] = {k: lat.download(k).astype(np.float64) for k in init}
            finally:
                lat.close()
    for (geo, op), o in out.items():
        ref = out[(0, 0, 1, 2), op]
        for k in init:
            assert rel(o[k], ref[k]) < 1e-6, (geo, op, k)


@pytest.mark.parametrize("zn_active", [1.0, 0.0])
@pytest.mark.parametrize("shape", [(6, 16, 250), (5, 12, 256), (4, 10, 128)])
def test_complex_shared_halo_rows(shape, zn_active):
    """Complex rows of two or four waves run two row groups per block that read their inner
    y-halo rows from each other's LDS staging (where ny divides into pairs of tiles): every
    row geometry gives identical BAOAB and overdamped fields, the flat tile agrees to float32
    rounding, and the default form matches the oracle; Z_N on, and off (the row kernels with
    the Z_N term compiled out)."""
    from tip.lattice import Lattice, KERNEL_FLAT, KERNEL_ROWS
    f, q, tab, prm = complex_setup(shape, sum(shape))
    init = {k: v.astype(np.float32) for k, v in
            dict(phi1=f[0], phi2=f[1], pi1=q[0], pi2=q[1]).items()}
    kw = dict(dt=prm["dt"], dx=prm["dx"], mu=prm["mu"], inv_a2=prm["inv_a2"],
              noise_scale=prm["ns"], seed=3, step0=0)
    out = {}
    geos = [(1, 0, 1, KERNEL_ROWS), (2, 3, 1, KERNEL_ROWS), (4, 0, 1, KERNEL_ROWS),
            (1, 0, 1, KERNEL_FLAT)]
    for geo in geos:
        if geo[3] == KERNEL_ROWS and shape[1] % geo[0]:
            continue
        for op in ("baoab", "overdamped"):
            lat = Lattice(shape, np.float32, complex_field=True)
            try:
                for k, v in init.items():
                    lat.upload(k, v)
                lat.set_table(*tab, odd=False)
                lat.set_zn(3.0, 1e8, zn_active)
                lat.set_geometry(*geo)
                if op == "baoab":
                    lat.step_baoab(4, gamma=prm["eta"], **kw)
                else:
                    lat.step_overdamped(4, eta_eff=prm["eta"], **kw)
                out[geo, op] = {k: lat.download(k) for k in init}
            finally:
                lat.close()
    for (geo, op), o in out.items():
        ref = out[(1, 0, 1, KERNEL_ROWS), op]
        for k in init:
            if geo[3] == KERNEL_FLAT:
                assert rel(o[k].astype(np.float64), ref[k].astype(np.float64)) < 1e-6, (geo, op)
            else:
                assert np.array_equal(o[k], ref[k]), (geo, op, k)
    # the default geometry (shared halos) through the drop-in against the oracle
    from tip import kernels_complex as KC
    a, b = (x.astype(np.float32) for x in f)
    f0 = [f[0].astype(np.float32).astype(np.float64), f[1].astype(np.float32).astype(np.float64)]
    q0 = [q[0].astype(np.float64), q[1].astype(np.float64)]
    for n in range(3):
        KC.overdamped_euler_step_table(a, b, prm["dt"], prm["dx"], prm["eta"], prm["mu"],
                                       prm["inv_a2"], tab[0], tab[1], tab[2], len(tab[0]),
                                       prm["ns"], O.noise_seed(n), zn=(3.0, 1e8, zn_active))
    (ra, rb), _ = oracle_complex("overdamped", f0, q0, tab, prm, (3.0, 1e8, zn_active), 3)
    assert rel(a.astype(np.float64), ra) < 2e-5 and rel(b.astype(np.float64), rb) < 2e-5


def test_row_tile_fallbacks_and_errors():
    """A shape the row tile cannot take falls back to the flat tile under the default geometry
    (complex nz > 512) and agrees with the oracle; requesting the row kernel for such a shape
    (or rows per thread not dividing ny) raises; a ragged real shape runs the row tile."""
    from tip.lattice import KERNEL_ROWS, Lattice
    f, q, tab, prm = complex_setup((3, 5, 600), 17)
    lat = Lattice((3, 5, 600), np.float32, complex_field=True)
    try:
        for k, v in zip(("phi1", "phi2", "pi1", "pi2"), (*f, *q)):
            lat.upload(k, v.astype(np.float32))
        lat.set_table(*tab)
        lat.step_overdamped(2, dt=prm["dt"], dx=prm["dx"], eta_eff=prm["eta"], mu=prm["mu"],
                            noise_scale=prm["ns"], seed=0, step0=0)
        got = lat.download("phi1").astype(np.float64)
        with pytest.raises(RuntimeError):
            lat.set_geometry(2, 0, True, KERNEL_ROWS)  # nz = 600 > 512
    finally:
        lat.close()
    a, b = (x.astype(np.float32).astype(np.float64) for x in f)
    for n in range(2):
        a, b, _, _ = O.langevin_complex(O.OVERDAMPED, a, b, None, None, prm["dt"], prm["dx"],
                                        prm["eta"], prm["mu"], 1.0, *tab,
                                        noise_seed=O.noise_seed(n), noise_scale=prm["ns"])
    assert rel(got, a) < 2e-5
    phi0, pi0, tab_r, prm_r = real_setup((6, 10, 64), 19)
    lat = Lattice((6, 10, 64), np.float32)
    try:
        lat.upload("phi", phi0.astype(np.float32))
        lat.upload("pi", pi0.astype(np.float32))
        lat.set_table(*tab_r, odd=True)
        with pytest.raises(RuntimeError):
            lat.set_geometry(4, 0, True, KERNEL_ROWS)  # 4 rows per thread do not divide ny = 10
        # the default geometry takes the row tile with one row per thread: the step runs
        lat.step_baoab(2, dt=prm_r["dt"], dx=prm_r["dx"], gamma=prm_r["eta"], mu=prm_r["mu"],
                       inv_a2=prm_r["inv_a2"], noise_scale=prm_r["ns"], seed=0, step0=0)
        got_phi = lat.download("phi").astype(np.float64)
    finally:
        lat.close()
    rp, _ = oracle_real("baoab", phi0.astype(np.float32).astype(np.float64),
                        pi0.astype(np.float32).astype(np.float64), tab_r, prm_r, 2)
    assert rel(got_phi, rp) < 2e-5
