"""The complex field at its own size (-m gpu): 256^3 float32, T0 = 6770 GeV, the complex
driver's couplings (V_p, lam = 1e-16), dx_phys = 1e-3, dt_phys = 2.5e-4, RK2.

Bars:
  * the fused complex step at 256^3 against the oracle's float32-mixed restatement of
    rk2_step_table (simulation/latticeSimComplex_numba.py:2067-2215) on the same hash noise,
    Z_N on: relative field error < 1e-5 after 3 steps (the float32 bar of the small-lattice
    complex test);
  * winding on evolved snapshots of a 256^3 driver run (Hubble expansion, the device table
    rebuilds): per-plaquette int8 windings, the winding buffer (the integers the reference's
    float64 sums round to) and the string count bit-exact
    against the oracle's compute_winding_number (:2683-2744) in float64 on the same float32
    snapshot (:4627-4629, :4661).
"""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

N = 256
SHAPE = (N, N, N)


def rel(a, b):
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))


def test_complex_256_steps_against_oracle():
    from tip import potential as P
    from tip.lattice import Lattice
    p = P.ModelParams.complex()
    T0, mu = 6770.0, 1000.0
    t, tmin, dinv = P.build_vprime_table(T0, 0.0, 1e5, 16384, p)
    dt, dx, eta = 0.25, 1.0, 1.0
    ns = P.noise_scale(eta, T0, dt, mu, 1e-3)
    zn = (3.0, 1e8, 1.0)
    rng = np.random.default_rng(256)
    th = 2 * np.pi * rng.random(SHAPE)
    r = 300.0 * np.abs(rng.normal(size=SHAPE))
    f = [(r * np.cos(th)).astype(np.float32), (r * np.sin(th)).astype(np.float32),
         np.zeros(SHAPE, np.float32), np.zeros(SHAPE, np.float32)]
    del th, r
    lat = Lattice(SHAPE, np.float32, complex_field=True)
    for k, v in zip(("phi1", "phi2", "pi1", "pi2"), f):
        lat.upload(k, v)
    lat.set_table(t, tmin, dinv)
    lat.set_zn(*zn)
    nsteps = 3
    lat.step(nsteps, dt=dt, dx=dx, eta_eff=eta, mu=mu, noise_scale=ns, seed=6)
    got = [lat.download(k) for k in ("phi1", "phi2", "pi1", "pi2")]
    lat.close()
    tmp = [np.empty(SHAPE, np.float32) for _ in range(4)]
    for n in range(nsteps):
        n1, n2 = O.noise_field(SHAPE, O.noise_seed(n, 6), ns, dtype=np.float32, pair=True)
        O.rk2_step_table_complex(*f, dt, dx, eta, mu, n1, n2, *tmp, 1.0, t, tmin, dinv, len(t),
                                 zn=zn)
    errs = [rel(g.astype(np.float64), a.astype(np.float64)) for g, a in zip(got, f)]
    print(f"complex 256^3, {nsteps} steps: float32 kernel vs oracle float32-mixed rel err "
          + " ".join(f"{e:.2e}" for e in errs), flush=True)
    assert max(errs) < 1e-5


def test_winding_bit_exact_on_evolved_256_snapshots():
    from tip import driver as D
    from tip.device import DeviceArray
    cfg = D.RunConfig(field="complex", Nx=N, Ny=N, Nz=N, Nt=1000, T0=6770.0, no_hubble=False,
                      integrator="rk2", report=500, seed=2, quiet=True, table_size=16384)
    run = D.Run(cfg)
    snaps = []
    observe = run.observe

    def snap(n, T):
        observe(n, T)
        plaq = DeviceArray((3,) + SHAPE, np.int8)
        wsum = DeviceArray(SHAPE, np.float32)
        strings, nplq = run.lat.winding_counts(plaq.ptr, wsum.ptr)
        snaps.append((n, run.lat.download("phi1"), run.lat.download("phi2"), plaq.download(),
                      wsum.download(), strings, nplq))
        plaq.free()
        wsum.free()

    run.observe = snap
    try:
        run.run()
    finally:
        run.close()
    assert [s[0] for s in snaps] == [500, 1000] and run.rebuilds >= 1
    for n, p1, p2, plaq, wsum, strings, nplq in snaps:
        w_ref, plaq_ref = O.winding(p1, p2)
        print(f"step {n}: {strings} string sites, {nplq} nonzero plaquettes, rho max "
              f"{float(np.hypot(p1, p2).max()):.3e}", flush=True)
        assert np.array_equal(plaq, plaq_ref), n
        # the reference's buffer holds the float64 sum of wrapped phase differences / 2 pi
        # (integers up to ~1e-18 residue); the kernel's is the exact integer sum
        assert np.array_equal(wsum, np.rint(w_ref).astype(np.float32)), n
        assert np.max(np.abs(w_ref - np.rint(w_ref))) < 1e-6
        assert strings == O.string_count(w_ref) and nplq == int(np.count_nonzero(plaq_ref))
        assert strings > 0
