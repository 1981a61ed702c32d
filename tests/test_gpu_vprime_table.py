"""The V' table producer on the device (tip_build_vprime_table, -m gpu) against the host
restatement tip.potential.build_vprime_table (pinned by the reference's run-log prints in
tests/test_potential.py).

Reference: Vprime_scalar + build_vprime_table, simulation/latticeSimeRescale_numba.py:805-849.
Bar: every entry within 1e-14 of the magnitude of its terms (|lam phi^3| + |mphi^2 phi| +
|thermal part|: the tree and thermal terms cancel near the barrier and the vacuum, so a
relative bar on V' itself would measure the cancellation, not the producer); the returned
(tmin, dinv) equal the host's exactly.  Cases: the real driver's model (V_correct x 20), the
complex driver's (V_p, no multiplicities), fermion-only, the lattice counterterm's spline,
ranges crossing 0, spanning the barrier and the spline clamp (u > 100), and 2 to 16384
points.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _scale(t, phi, p):
    tree = p.lam * phi ** 3 - p.mphi ** 2 * phi
    return np.abs(p.lam * phi ** 3) + np.abs(p.mphi ** 2 * phi) + np.abs(t - tree)


CASES = [
    ("real", 7000.0, 0.0, 1.5e15, 8192, 0.0),
    ("real", 7000.0, -2e4, 2e4, 16384, 0.0),
    ("real", 6770.0, -3.3e5, 1.1e7, 4097, 0.0),
    ("complex", 6770.0, 0.0, 1e5, 16384, 0.0),
    ("complex", 7350.0, 0.0, 3e3, 2, 0.0),
    ("fermion_only", 4912.0, -1e8, 1e8, 1001, 0.0),
    ("real", 7000.0, 0.0, 4e4, 2049, 0.002),  # counterterm spline (dx_phys = 0.002)
]


@pytest.mark.parametrize("case", CASES)
def test_device_table_equals_host_producer(case):
    from tip import potential as P
    from tip import vprime_device as VD
    kind, T, lo, hi, n, ct_dx = case
    p = {"real": P.ModelParams.real(), "complex": P.ModelParams.complex(),
         "fermion_only": P.ModelParams.real("fermion_only")}[kind]
    sp = P.dj_spline(x_cut=P.counterterm_xcut(ct_dx, T)) if ct_dx else P.dj_spline()
    want, tmin, dinv = P.build_vprime_table(T, lo, hi, n, p, sp)
    dev, tmin_d, dinv_d = VD.build_vprime_table(T, lo, hi, n, p, sp)
    got = dev.download()
    dev.free()
    assert tmin_d == tmin and dinv_d == dinv
    phi = np.linspace(lo, hi, n)
    err = np.abs(got - want) / np.maximum(_scale(want, phi, p), 1e-300)
    print(f"{kind} T={T} [{lo:.3g}, {hi:.3g}] n={n}: max err / term scale {err.max():.2e}, "
          f"bitwise equal {np.mean(got == want):.3f}", flush=True)
    assert err.max() <= 1e-14


def test_lattice_build_table_steps_like_set_table():
    """A lattice whose table was built on the device steps as one given the host's table
    (float64 fields: within 1e-13 relative after 5 steps; the tables differ by ulps)."""
    from tip import potential as P
    from tip.lattice import Lattice
    p = P.ModelParams.real()
    shape = (16, 16, 64)
    rng = np.random.default_rng(4)
    phi0 = rng.normal(size=shape) * 3e3
    out = []
    for dev in (True, False):
        lat = Lattice(shape, np.float64)
        lat.upload("phi", phi0)
        lat.upload("pi", np.zeros(shape))
        if dev:
            lat.build_table(7000.0, -2e4, 2e4, 4096, p)
        else:
            lat.set_table(*P.build_vprime_table(7000.0, -2e4, 2e4, 4096, p))
        lat.step(5, dt=0.25, dx=2.0, eta_eff=7.0, mu=1000.0, noise_scale=1750.0, seed=2)
        out.append(lat.download("phi"))
        lat.close()
    assert np.linalg.norm(out[0] - out[1]) <= 1e-13 * np.linalg.norm(out[1])
