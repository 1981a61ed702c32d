"""Pins of the oracle's overdamped Euler-Maruyama and BAOAB stages (CPU).

Reference: overdamped_euler_step_table and baoab_step_table,
simulation/latticeSimeRescale_numba.py:1974-2192 (real) and
simulation/latticeSimComplex_numba.py:2223-2517 (complex), in the race-free (Jacobi) form.

Each check is chosen so that a plausible restatement mistake fails it:
  * a Fourier mode with a linear V' table and no noise: the overdamped step multiplies it by
    1 - (dt/eta)(inv_a2 lam + c/mu^2); a BAOAB step applies the product of the five linear maps
    B A O A B (a wrong half step, a kick with the wrong sign, an O step on the wrong side of a
    drift or a wrong c1 changes the matrix)
  * noise members: on a zero field with a zero table the overdamped step is exactly ns z and
    the BAOAB open stage gives pi = c2 z with c2 = sqrt(1 - exp(-2 gamma dt)) ns (cos member
    for the real field, (cos, sin) for the complex field, the RK2 path's site seeds)
  * the fused stage: open, fused x (k-1), close is bit-identical to k open+close steps
  * a literal loop restatement of the per-site text on a tiny ragged lattice (bitwise)
  * the Ornstein-Uhlenbeck stationary variance ns^2 (fluctuation-dissipation)
  * complex: U(1) covariance with Z_N off, and the uniform field reduces to the real ODE
"""
import math

import numpy as np
import pytest

from oracle import oracle as O


def linear_table(c, lo=-1e3, hi=1e3, n=65):
    x = np.linspace(lo, hi, n)
    return c * x, lo, 1.0 / ((hi - lo) / (n - 1))


def mode_field(shape, m):
    i, j, k = np.meshgrid(*[np.arange(n) for n in shape], indexing="ij")
    return np.cos(2 * np.pi * (m[0] * i / shape[0] + m[1] * j / shape[1] + m[2] * k / shape[2]))


def lam(shape, m, dx):
    return (4.0 / dx ** 2) * sum(math.sin(math.pi * mm / n) ** 2 for mm, n in zip(m, shape))


def baoab_matrix(dt, gamma, w2):
    h = 0.5 * dt
    B = np.array([[1.0, 0.0], [-h * w2, 1.0]])
    A = np.array([[1.0, h], [0.0, 1.0]])
    Om = np.array([[1.0, 0.0], [0.0, math.exp(-gamma * dt)]])
    return B @ A @ Om @ A @ B


SHAPE, MODE = (8, 6, 10), (1, 2, 3)
DX, DT, ETA, MU, INV_A2, C = 2.0, 0.25, 7.0, 1000.0, 0.9, 3.0e5


def test_overdamped_fourier_mode():
    mode = mode_field(SHAPE, MODE)
    phi = 5.0 * mode
    tab, tmin, dinv = linear_table(C)
    g = 1.0 - (DT / ETA) * (INV_A2 * lam(SHAPE, MODE, DX) + C / MU ** 2)
    for _ in range(15):
        phi, _ = O.langevin(O.OVERDAMPED, phi, None, DT, DX, ETA, MU, INV_A2, tab, tmin, dinv)
    np.testing.assert_allclose(phi, 5.0 * g ** 15 * mode, rtol=0, atol=1e-12 * 5.0)


@pytest.mark.parametrize("fused", [False, True])
def test_baoab_fourier_mode(fused):
    mode = mode_field(SHAPE, MODE)
    phi, pi = 5.0 * mode, -2.0 * mode
    tab, tmin, dinv = linear_table(C)
    M = baoab_matrix(DT, ETA, INV_A2 * lam(SHAPE, MODE, DX) + C / MU ** 2)
    y = np.array([5.0, -2.0])
    n = 12
    for _ in range(n):
        y = M @ y
    if fused:
        phi, pi = O.langevin(O.BAOAB_OPEN, phi, pi, DT, DX, ETA, MU, INV_A2, tab, tmin, dinv)
        for _ in range(n - 1):
            phi, pi = O.langevin(O.BAOAB_FUSED, phi, pi, DT, DX, ETA, MU, INV_A2, tab, tmin,
                                 dinv)
        _, pi = O.langevin(O.BAOAB_CLOSE, phi, pi, DT, DX, ETA, MU, INV_A2, tab, tmin, dinv)
    else:
        for _ in range(n):
            phi, pi = O.baoab_step(phi, pi, DT, DX, ETA, MU, INV_A2, tab, tmin, dinv)
    np.testing.assert_allclose(phi, y[0] * mode, rtol=0, atol=1e-12 * abs(y).max())
    np.testing.assert_allclose(pi, y[1] * mode, rtol=0, atol=1e-12 * abs(y).max())


def test_noise_members_real():
    shape, ns, seed = (4, 5, 33), 0.37, 12345
    zero = np.zeros(shape)
    tab = np.zeros(9)
    phi, _ = O.langevin(O.OVERDAMPED, zero, None, DT, DX, ETA, MU, 1.0, tab, 0.0, 1.0,
                        noise_seed=seed, noise_scale=ns)
    # (libm's cos and the sincos pair the noise-field routine uses may differ in the last bit)
    np.testing.assert_allclose(phi, O.noise_field(shape, seed, ns), rtol=1e-15, atol=0)
    gamma = 3.0
    c2 = math.sqrt(1.0 - math.exp(-2.0 * gamma * DT)) * ns
    phi, pi = O.langevin(O.BAOAB_OPEN, zero, zero, DT, DX, gamma, MU, 1.0, tab, 0.0, 1.0,
                         noise_seed=seed, noise_scale=ns)
    z = O.noise_field(shape, seed, 1.0)
    np.testing.assert_allclose(pi, c2 * z, rtol=1e-15, atol=0)
    np.testing.assert_allclose(phi, 0.5 * DT * c2 * z, rtol=1e-15, atol=0)


def test_noise_members_complex():
    shape, ns, seed = (3, 4, 21), 0.5, 77
    zero = np.zeros(shape)
    tab = np.zeros(9)
    a, b, _, _ = O.langevin_complex(O.OVERDAMPED, zero, zero, None, None, DT, DX, ETA, MU, 1.0,
                                    tab, 0.0, 1.0, noise_seed=seed, noise_scale=ns)
    z1, z2 = O.noise_field(shape, seed, ns, pair=True)
    np.testing.assert_allclose(a, z1, rtol=1e-15, atol=0)
    np.testing.assert_allclose(b, z2, rtol=1e-15, atol=0)


def test_fused_chain_bit_identical(rng):
    shape = (6, 5, 7)
    phi = rng.normal(size=shape) * 300.0
    pi = rng.normal(size=shape) * 100.0
    tab = rng.normal(size=33) * 1e5
    tmin, dinv = -2000.0, 32 / 4000.0
    kw = dict(dt=DT, dx=DX, eta_or_gamma=ETA, mu=MU, inv_a2=0.8, table=tab, tmin=tmin,
              dinv=dinv)
    a, b = phi, pi
    for n in range(4):
        a, b = O.baoab_step(a, b, DT, DX, ETA, MU, 0.8, tab, tmin, dinv, noise_seed=O.noise_seed(n),
                            noise_scale=2.0)
    c, d = O.langevin(O.BAOAB_OPEN, phi, pi, noise_seed=O.noise_seed(0), noise_scale=2.0, **kw)
    for n in range(1, 4):
        c, d = O.langevin(O.BAOAB_FUSED, c, d, noise_seed=O.noise_seed(n), noise_scale=2.0, **kw)
    _, d = O.langevin(O.BAOAB_CLOSE, c, d, **kw)
    assert np.array_equal(a, c) and np.array_equal(b, d)


# ----------------------------------------------------------------------------------------
# literal loop restatement of the per-site text (latticeSimeRescale_numba.py:2035-2060,
# 2073, 2092-2112, 2170-2192) in the Jacobi form
# ----------------------------------------------------------------------------------------
def _lookup_py(x, t, tmin, dinv):
    fidx = (x - tmin) * dinv
    idx = int(fidx)
    idx = min(max(idx, 0), len(t) - 2)
    frac = fidx - idx
    return t[idx] + frac * (t[idx + 1] - t[idx])


def _force_py(phi, i, j, k, inv_dx2, inv_mu2, inv_a2, t, tmin, dinv):
    nx, ny, nz = phi.shape
    ph = phi[i, j, k]
    lap = (phi[(i + 1) % nx, j, k] + phi[(i - 1) % nx, j, k] + phi[i, (j + 1) % ny, k]
           + phi[i, (j - 1) % ny, k] + phi[i, j, (k + 1) % nz] + phi[i, j, (k - 1) % nz]
           - 6.0 * ph) * inv_dx2
    dV = _lookup_py(ph, t, tmin, dinv)
    return inv_a2 * lap - dV * inv_mu2


def _z(site, seed):
    x1, x2, u1, u2 = O.hash_pair_py((site * 73856093 % 2 ** 64) ^ seed)
    return math.sqrt(-2.0 * math.log(u1)) * math.cos(2.0 * math.pi * u2)


def test_loop_restatement(rng):
    shape = (3, 4, 5)
    phi = rng.normal(size=shape) * 300.0
    pi = rng.normal(size=shape) * 100.0
    t = rng.normal(size=17) * 1e5
    tmin, dinv = -1000.0, 16 / 2000.0
    dt, dx, gamma, mu, inv_a2, ns, seed = 0.3, 1.7, 2.5, 900.0, 0.85, 1.3, 99
    inv_dx2, inv_mu2, half_dt = 1.0 / (dx * dx), 1.0 / (mu * mu), 0.5 * dt
    c1 = math.exp(-gamma * dt)
    c2 = math.sqrt(1.0 - c1 * c1) * ns
    od = np.empty(shape)
    po, qo = np.empty(shape), np.empty(shape)
    for i in range(shape[0]):
        for j in range(shape[1]):
            for k in range(shape[2]):
                s = (i * shape[1] + j) * shape[2] + k
                F = _force_py(phi, i, j, k, inv_dx2, inv_mu2, inv_a2, t, tmin, dinv)
                z = _z(s, seed)
                od[i, j, k] = phi[i, j, k] + (dt / gamma) * F + ns * z
                pi_v = pi[i, j, k] + half_dt * F
                ph = phi[i, j, k] + half_dt * pi_v
                pi_v = c1 * pi_v + c2 * z
                po[i, j, k] = ph + half_dt * pi_v
                qo[i, j, k] = pi_v
    got, _ = O.langevin(O.OVERDAMPED, phi, None, dt, dx, gamma, mu, inv_a2, t, tmin, dinv,
                        noise_seed=seed, noise_scale=ns)
    assert np.array_equal(got, od)
    a, b = O.langevin(O.BAOAB_OPEN, phi, pi, dt, dx, gamma, mu, inv_a2, t, tmin, dinv,
                      noise_seed=seed, noise_scale=ns)
    assert np.array_equal(a, po) and np.array_equal(b, qo)
    close = np.empty(shape)
    for i in range(shape[0]):
        for j in range(shape[1]):
            for k in range(shape[2]):
                F = _force_py(po, i, j, k, inv_dx2, inv_mu2, inv_a2, t, tmin, dinv)
                close[i, j, k] = qo[i, j, k] + half_dt * F
    _, d = O.langevin(O.BAOAB_CLOSE, a, b, dt, dx, gamma, mu, inv_a2, t, tmin, dinv)
    assert np.array_equal(d, close)


def test_ou_stationary_variance():
    """With the force switched off (zero table, dx so large the Laplacian vanishes) the BAOAB
    momentum is an Ornstein-Uhlenbeck process whose stationary variance is ns^2."""
    shape, ns, gamma, dt = (32, 32, 32), 0.8, 0.5, 0.1
    phi = np.zeros(shape)
    pi = np.zeros(shape)
    tab = np.zeros(5)
    for n in range(120):  # c1^(2n) = exp(-12)
        phi, pi = O.langevin(O.BAOAB_FUSED if n else O.BAOAB_OPEN, phi, pi, dt, 1e150, gamma,
                             1.0, 1.0, tab, 0.0, 1.0, noise_seed=O.noise_seed(n),
                             noise_scale=ns)
    assert abs(pi.var() / ns ** 2 - 1.0) < 0.03
    assert abs(pi.mean()) < 0.03 * ns


def test_complex_u1_covariance(rng):
    shape = (5, 4, 6)
    r = 200.0 + 50.0 * rng.random(size=shape)
    th = rng.uniform(-np.pi, np.pi, size=shape)
    p1, p2 = r * np.cos(th), r * np.sin(th)
    q1, q2 = rng.normal(size=shape) * 10.0, rng.normal(size=shape) * 10.0
    x = np.linspace(0.0, 400.0, 129)
    tab, tmin, dinv = 3.0e5 * x - 2e-3 * x ** 3, 0.0, 128 / 400.0
    alpha = 0.7
    ca, sa = math.cos(alpha), math.sin(alpha)
    rot = lambda a, b: (ca * a - sa * b, sa * a + ca * b)
    for op in (O.OVERDAMPED, O.BAOAB_OPEN, O.BAOAB_CLOSE):
        out = O.langevin_complex(op, p1, p2, q1, q2, DT, DX, ETA, MU, 0.9, tab, tmin, dinv)
        r1, r2 = rot(p1, p2)
        s1, s2 = rot(q1, q2)
        out_r = O.langevin_complex(op, r1, r2, s1, s2, DT, DX, ETA, MU, 0.9, tab, tmin, dinv)
        for pair in ((0, 1), (2, 3)):
            if out[pair[0]] is None:
                continue
            e1, e2 = rot(out[pair[0]], out[pair[1]])
            scale = np.abs(out[pair[0]]).max()
            np.testing.assert_allclose(out_r[pair[0]], e1, rtol=0, atol=1e-11 * scale)
            np.testing.assert_allclose(out_r[pair[1]], e2, rtol=0, atol=1e-11 * scale)


def test_complex_uniform_reduces_to_real():
    """A uniform complex field on the real axis with a radial table V'(rho) = c rho evolves
    exactly as the real field with V'(phi) = c phi."""
    shape = (4, 3, 5)
    tab_r, tmin, dinv = linear_table(C, lo=0.0, hi=2e3)
    phi = np.full(shape, 30.0)
    pi = np.full(shape, -4.0)
    zero = np.zeros(shape)
    a, b = O.baoab_step(phi, pi, DT, DX, ETA, MU, 1.0, tab_r, tmin, dinv)
    c1, c2, d1, d2 = O.langevin_complex(O.BAOAB_OPEN, phi, zero, pi, zero, DT, DX, ETA, MU, 1.0,
                                        tab_r, tmin, dinv)
    _, _, d1, d2 = O.langevin_complex(O.BAOAB_CLOSE, c1, c2, d1, d2, DT, DX, ETA, MU, 1.0, tab_r,
                                      tmin, dinv)
    assert np.array_equal(c1, a) and np.array_equal(d1, b)
    assert not c2.any() and not d2.any()
