"""Pins of the oracle's RK2 step, lookup and complex force (CPU).

Each check is chosen so that a plausible restatement mistake fails it:
  * lookup hand values -- truncation toward zero, index clamp, unclamped extrapolation, odd table
  * a Fourier mode with a linear V' table and no noise evolves by the exact RK2 matrix
    M = I + dt A + dt^2 A^2 / 2, A = [[0, 1], [-w^2, -eta]], w^2 = inv_a2 lam_m + c/mu^2
    (wrong sign, dropped term, wrong stage argument or wrong Laplacian all change M)
  * translation covariance under periodic shifts (wrong wrap index)
  * brute force: an independent loop restatement of the reference text on tiny ragged
    lattices (float64 bitwise; float32-mixed bitwise), real and complex with Z_N
  * U(1) phase covariance of the complex step with Z_N off
"""
import math

import numpy as np
import pytest

from oracle import oracle as O


# ----------------------------------------------------------------------------------------
# lookup (latticeSimeRescale_numba.py:1923-1930)
# ----------------------------------------------------------------------------------------
def test_lookup_hand_values():
    t = np.array([10.0, 20.0, 40.0, 80.0])
    x = np.array([0.5, 2.5, 3.5, -0.5, -2.5, 0.0, 3.0, 1.0])
    got = O.lookup(x, t, 0.0, 1.0)
    # -0.5: int() truncates to 0, frac -0.5 -> 5; -2.5: int -2 -> clamp 0, frac -2.5 -> -15
    # 3.5: clamp idx 2, frac 1.5 -> 100; 3.0: idx 3 -> clamp 2, frac 1 -> 80
    np.testing.assert_array_equal(got, [15.0, 60.0, 100.0, 5.0, -15.0, 10.0, 80.0, 20.0])
    odd = O.lookup(np.array([-0.5, 0.5, -3.5]), t, 0.0, 1.0, odd=True)
    np.testing.assert_array_equal(odd, [-15.0, 15.0, -100.0])
    # non-unit spacing and origin
    got = O.lookup(np.array([1.25]), t, 1.0, 2.0)  # fidx = 0.5
    assert got[0] == 15.0


# ----------------------------------------------------------------------------------------
# closed form: one Fourier mode, linear table, zero noise
# ----------------------------------------------------------------------------------------
def linear_table(c, lo=-1e3, hi=1e3, n=65):
    x = np.linspace(lo, hi, n)
    return c * x, lo, 1.0 / ((hi - lo) / (n - 1))


def mode_matrix(dt, eta, w2):
    A = np.array([[0.0, 1.0], [-w2, -eta]])
    return np.eye(2) + dt * A + 0.5 * dt * dt * A @ A


@pytest.mark.parametrize("dtype", [np.float64])
def test_fourier_mode_exact_matrix(dtype):
    shape = (8, 6, 10)
    m = (1, 2, 3)
    dx, dt, eta, mu, inv_a2, c = 2.0, 0.25, 7.0, 1000.0, 0.9, 3.0e5
    i, j, k = np.meshgrid(*[np.arange(n) for n in shape], indexing="ij")
    mode = np.cos(2 * np.pi * (m[0] * i / shape[0] + m[1] * j / shape[1] + m[2] * k / shape[2]))
    phi = (5.0 * mode).astype(dtype)
    pi = (-2.0 * mode).astype(dtype)
    tab, tmin, dinv = linear_table(c)
    lam_m = (4.0 / dx ** 2) * sum(math.sin(math.pi * mm / n) ** 2 for mm, n in zip(m, shape))
    w2 = inv_a2 * lam_m + c / mu ** 2
    M = mode_matrix(dt, eta, w2)
    nsteps = 20
    y = np.array([5.0, -2.0])
    for _ in range(nsteps):
        y = M @ y
    pt, qt = np.empty_like(phi), np.empty_like(pi)
    for _ in range(nsteps):
        O.rk2_step_table(phi, pi, dt, dx, eta, mu, None, pt, qt, inv_a2, tab, tmin, dinv, len(tab))
    np.testing.assert_allclose(phi, y[0] * mode, rtol=0, atol=1e-12 * abs(y).max())
    np.testing.assert_allclose(pi, y[1] * mode, rtol=0, atol=1e-12 * abs(y).max())


def test_uniform_field_is_the_ode():
    shape = (5, 7, 9)
    dt, eta, mu, c = 0.25, 7.0, 1000.0, -1.0e6
    phi = np.full(shape, 3.0)
    pi = np.full(shape, 1.5)
    tab, tmin, dinv = linear_table(c)
    M = mode_matrix(dt, eta, c / mu ** 2)
    y = np.array([3.0, 1.5])
    pt, qt = np.empty_like(phi), np.empty_like(pi)
    for _ in range(10):
        O.rk2_step_table(phi, pi, dt, 2.0, eta, mu, None, pt, qt, 1.0, tab, tmin, dinv, len(tab))
        y = M @ y
    np.testing.assert_allclose(phi, y[0], rtol=1e-13)
    np.testing.assert_allclose(pi, y[1], rtol=1e-13)


# ----------------------------------------------------------------------------------------
# translation covariance
# ----------------------------------------------------------------------------------------
def test_translation_covariance(rng):
    shape = (6, 5, 7)
    phi = rng.normal(size=shape) * 300.0
    pi = rng.normal(size=shape) * 100.0
    noise = rng.normal(size=shape) * 50.0
    tab = rng.normal(size=33) * 1e5
    tmin, dinv = -2000.0, 32 / 4000.0
    shift = (2, 3, 4)

    def step(a, b, n):
        a, b = a.copy(), b.copy()
        O.rk2_step_table(a, b, 0.25, 2.0, 7.0, 1000.0, n, np.empty_like(a), np.empty_like(b),
                         0.8, tab, tmin, dinv, len(tab))
        return a, b

    r = lambda x: np.roll(x, shift, axis=(0, 1, 2))
    a1, b1 = step(r(phi), r(pi), r(noise))
    a2, b2 = step(phi, pi, noise)
    assert np.array_equal(a1, r(a2)) and np.array_equal(b1, r(b2))


# ----------------------------------------------------------------------------------------
# brute force: an independent loop restatement of the reference text
# ----------------------------------------------------------------------------------------
def _lookup_py(x, t, tmin, dinv):
    fidx = (x - tmin) * dinv
    idx = int(fidx)
    idx = min(max(idx, 0), len(t) - 2)
    frac = fidx - idx
    return t[idx] + frac * (t[idx + 1] - t[idx])


def _step_py(phi, pi, dt, dx, eta, mu, noise, inv_a2, t, tmin, dinv, f32=False):
    nx, ny, nz = phi.shape
    inv_dx2, inv_mu2, h = 1.0 / (dx * dx), 1.0 / (mu * mu), 0.5 * dt
    cast = (lambda v: float(np.float32(v))) if f32 else float
    def nsum(a, i, j, k):
        ip, im = (i + 1) % nx, (i - 1) % nx
        jp, jm = (j + 1) % ny, (j - 1) % ny
        kp, km = (k + 1) % nz, (k - 1) % nz
        vals = [a[ip, j, k], a[im, j, k], a[i, jp, k], a[i, jm, k], a[i, j, kp], a[i, j, km]]
        s = vals[0]
        for v in vals[1:]:
            s = s + v  # float32 + float32 stays float32 (numpy scalars), as Numba types it
        return float(s)
    pt = np.empty_like(phi)
    qt = np.empty_like(pi)
    for i in range(nx):
        for j in range(ny):
            for k in range(nz):
                ph, pv = float(phi[i, j, k]), float(pi[i, j, k])
                lap = (nsum(phi, i, j, k) - 6.0 * ph) * inv_dx2
                kpi = inv_a2 * lap - eta * pv - _lookup_py(ph, t, tmin, dinv) * inv_mu2
                pt[i, j, k] = ph + h * pv
                qt[i, j, k] = pv + h * kpi
    out_phi, out_pi = phi.copy(), pi.copy()
    for i in range(nx):
        for j in range(ny):
            for k in range(nz):
                ph, pv = float(pt[i, j, k]), float(qt[i, j, k])
                lap = (nsum(pt, i, j, k) - 6.0 * ph) * inv_dx2
                kpi = inv_a2 * lap - eta * pv - _lookup_py(ph, t, tmin, dinv) * inv_mu2
                out_phi[i, j, k] = cast(float(phi[i, j, k]) + dt * pv)
                out_pi[i, j, k] = cast(float(pi[i, j, k]) + (dt * kpi + float(noise[i, j, k])))
    return out_phi, out_pi


@pytest.mark.parametrize("shape", [(3, 4, 5), (1, 2, 3), (4, 1, 6)])
def test_brute_force_f64(rng, shape):
    phi = rng.normal(size=shape) * 800.0
    pi = rng.normal(size=shape) * 200.0
    noise = rng.normal(size=shape) * 30.0
    t = rng.normal(size=17) * 1e5
    tmin, dinv = -1500.0, 16 / 3000.0
    ep, eq = _step_py(phi, pi, 0.25, 2.0, 7.0, 1000.0, noise, 0.95, t, tmin, dinv)
    O.rk2_step_table(phi, pi, 0.25, 2.0, 7.0, 1000.0, noise, np.empty_like(phi),
                     np.empty_like(pi), 0.95, t, tmin, dinv, len(t))
    assert np.array_equal(phi, ep) and np.array_equal(pi, eq)


def test_brute_force_f32mixed(rng):
    shape = (3, 4, 5)
    phi = (rng.normal(size=shape) * 800.0).astype(np.float32)
    pi = (rng.normal(size=shape) * 200.0).astype(np.float32)
    noise = (rng.normal(size=shape) * 30.0).astype(np.float32)
    t = rng.normal(size=17) * 1e5
    tmin, dinv = -1500.0, 16 / 3000.0
    ep, eq = _step_py(phi, pi, 0.25, 2.0, 7.0, 1000.0, noise, 0.95, t, tmin, dinv, f32=True)
    O.rk2_step_table(phi, pi, 0.25, 2.0, 7.0, 1000.0, noise, np.empty_like(phi),
                     np.empty_like(pi), 0.95, t, tmin, dinv, len(t))
    assert np.array_equal(phi, ep.astype(np.float32)) and np.array_equal(pi, eq.astype(np.float32))


# ----------------------------------------------------------------------------------------
# complex field
# ----------------------------------------------------------------------------------------
def _complex_step_py(f1, f2, p1, p2, dt, dx, eta, mu, n1, n2, inv_a2, t, tmin, dinv, zn):
    nx, ny, nz = f1.shape
    inv_dx2, inv_mu2, h = 1.0 / (dx * dx), 1.0 / (mu * mu), 0.5 * dt
    zn_n, zn_dv, zn_act = zn

    def lap(a, i, j, k):
        s = (a[(i + 1) % nx, j, k] + a[(i - 1) % nx, j, k] + a[i, (j + 1) % ny, k]
             + a[i, (j - 1) % ny, k] + a[i, j, (k + 1) % nz] + a[i, j, (k - 1) % nz])
        return (s - 6.0 * a[i, j, k]) * inv_dx2

    def force(a, b):
        rho = math.sqrt(a * a + b * b)
        rs = max(rho, 1e-30)
        d = _lookup_py(rho, t, tmin, dinv)
        d1, d2 = d * a / rs, d * b / rs
        if zn_act > 0.5 and zn_n > 0.5 and rho > 1e-20:
            sn = math.sin(zn_n * math.atan2(b, a))
            ir2 = (1.0 / rs) * (1.0 / rs)
            d1 += -zn_n * zn_dv * sn * b * ir2
            d2 += zn_n * zn_dv * sn * a * ir2
        return d1, d2

    def stage(a, b, u, v):
        ka, kb = np.empty_like(a), np.empty_like(b)
        for i in range(nx):
            for j in range(ny):
                for k in range(nz):
                    d1, d2 = force(a[i, j, k], b[i, j, k])
                    ka[i, j, k] = inv_a2 * lap(a, i, j, k) - eta * u[i, j, k] - d1 * inv_mu2
                    kb[i, j, k] = inv_a2 * lap(b, i, j, k) - eta * v[i, j, k] - d2 * inv_mu2
        return ka, kb

    ka, kb = stage(f1, f2, p1, p2)
    a, b = f1 + h * p1, f2 + h * p2
    u, v = p1 + h * ka, p2 + h * kb
    ka2, kb2 = stage(a, b, u, v)
    return f1 + dt * u, f2 + dt * v, p1 + (dt * ka2 + n1), p2 + (dt * kb2 + n2)


@pytest.mark.parametrize("zn", [(0.0, 0.0, 0.0), (3.0, 2.5e9, 1.0)])
def test_complex_brute_force_f64(rng, zn):
    shape = (3, 4, 5)
    f1, f2 = rng.normal(size=shape) * 900.0, rng.normal(size=shape) * 900.0
    p1, p2 = rng.normal(size=shape) * 200.0, rng.normal(size=shape) * 200.0
    n1, n2 = rng.normal(size=shape) * 40.0, rng.normal(size=shape) * 40.0
    t = np.abs(rng.normal(size=21)) * 1e5
    tmin, dinv = 0.0, 20 / 4000.0
    exp = _complex_step_py(f1, f2, p1, p2, 0.25, 1.0, 1.0, 1000.0, n1, n2, 0.97, t, tmin, dinv, zn)
    got = [f1.copy(), f2.copy(), p1.copy(), p2.copy()]
    tmp = [np.empty(shape) for _ in range(4)]
    O.rk2_step_table_complex(*got, 0.25, 1.0, 1.0, 1000.0, n1, n2, *tmp, 0.97, t, tmin, dinv,
                             len(t), zn=zn)
    for g, e in zip(got, exp):
        np.testing.assert_allclose(g, e, rtol=1e-13, atol=1e-9)


def test_complex_phase_covariance(rng):
    shape = (4, 4, 6)
    f1, f2 = rng.normal(size=shape) * 900.0, rng.normal(size=shape) * 900.0
    p1, p2 = rng.normal(size=shape) * 200.0, rng.normal(size=shape) * 200.0
    n1, n2 = rng.normal(size=shape) * 40.0, rng.normal(size=shape) * 40.0
    t = np.abs(rng.normal(size=21)) * 1e5
    a = 0.7
    ca, sa = math.cos(a), math.sin(a)
    rot = lambda x, y: (ca * x - sa * y, sa * x + ca * y)

    def step(f1, f2, p1, p2, n1, n2):
        got = [f1.copy(), f2.copy(), p1.copy(), p2.copy()]
        tmp = [np.empty(shape) for _ in range(4)]
        O.rk2_step_table_complex(*got, 0.25, 1.0, 1.0, 1000.0, n1, n2, *tmp, 1.0, t, 0.0,
                                 20 / 4000.0, len(t))
        return got

    r1 = step(*rot(f1, f2), *rot(p1, p2), *rot(n1, n2))
    r0 = step(f1, f2, p1, p2, n1, n2)
    e1, e2 = rot(r0[0], r0[1])
    e3, e4 = rot(r0[2], r0[3])
    for g, e in zip(r1, (e1, e2, e3, e4)):
        np.testing.assert_allclose(g, e, rtol=1e-11, atol=1e-8)


def test_f32mixed_tracks_f64(rng):
    shape = (8, 8, 8)
    phi = rng.normal(size=shape) * 500.0
    pi = rng.normal(size=shape) * 100.0
    t, tmin, dinv = linear_table(-1.0e6, -1e5, 1e5, 257)
    params = dict(dt=0.25, dx=2.0, eta_eff=7.0, mu=1000.0, inv_a2=1.0, noise_scale=50.0)
    a, b = O.run_real(phi, pi, 20, params, t, tmin, dinv)
    c, d = O.run_real(phi.astype(np.float32), pi.astype(np.float32), 20, params, t, tmin, dinv)
    assert np.linalg.norm(c - a) / np.linalg.norm(a) < 1e-5
    assert np.linalg.norm(d - b) / np.linalg.norm(b) < 1e-5
