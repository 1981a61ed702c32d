"""Fluctuation-dissipation checks of the three integrators on the GPU (-m gpu), with the
acceptance bars of the reference's own Langevin validation (tools/validate_langevin.py:
1-30, 148-220): equipartition <pi^2> -> T / (mu^2 dx_phys^3) within 10 %, <pi^2>/target
independent of T within 0.12 (linearity), and a white velocity spectrum |pi_k|^2 (relative
spread < 0.25 in the sub-Nyquist band).

A massive free field (V' = c phi, linear table) at fixed T is run to equilibrium; because the
field is linear, the stationary covariance of every Fourier mode under each discrete scheme
is known exactly (a 2 x 2 Lyapunov equation per mode: C = M C M^T + b b^T with the scheme's
one-step matrix M and noise vector b; the overdamped scheme is scalar), so the measured
variance is also held to that exact value within 1 % -- a check of the noise amplitude, its
placement inside the step and the scheme's own damping, not only of the continuum limit.
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MU, DX_PHYS, DT_PHYS, ETA_PHYS, M2 = 1000.0, 1e-3, 2.5e-5, 1000.0, 0.5
N = 64


def lam_grid(n, dx):
    m = np.arange(n)
    s = (4.0 / dx ** 2) * np.sin(np.pi * m / n) ** 2
    return (s[:, None, None] + s[None, :, None] + s[None, None, :]).ravel()


def stationary_pi_var(scheme, lam, dt, eta, ns):
    """Mode-averaged stationary <pi^2> of the linear scheme (exact Lyapunov solution)."""
    w2 = lam + M2
    one = np.ones_like(w2)
    zero = np.zeros_like(w2)
    if scheme == "rk2":
        # M = I + dt A + dt^2 A^2 / 2, A = [[0, 1], [-w2, -eta]]; noise added to pi at the end
        a11, a12, a21, a22 = zero, one, -w2, -eta * one
        sq11 = a11 * a11 + a12 * a21
        sq12 = a11 * a12 + a12 * a22
        sq21 = a21 * a11 + a22 * a21
        sq22 = a21 * a12 + a22 * a22
        m11 = 1 + dt * a11 + 0.5 * dt * dt * sq11
        m12 = dt * a12 + 0.5 * dt * dt * sq12
        m21 = dt * a21 + 0.5 * dt * dt * sq21
        m22 = 1 + dt * a22 + 0.5 * dt * dt * sq22
        b1, b2 = zero, ns * one
    else:
        # BAOAB: M = B A O A B; the O-step noise c2 z passes through the last A and B
        h = 0.5 * dt
        c1 = math.exp(-eta * dt)
        c2 = math.sqrt(1.0 - c1 * c1) * ns

        def mat(a, b, c, d):
            return [a, b, c, d]

        def mul(P, Q):
            return [P[0] * Q[0] + P[1] * Q[2], P[0] * Q[1] + P[1] * Q[3],
                    P[2] * Q[0] + P[3] * Q[2], P[2] * Q[1] + P[3] * Q[3]]
        Bm = mat(one, zero, -h * w2, one)
        Am = mat(one, h * one, zero, one)
        Om = mat(one, zero, zero, c1 * one)
        M = mul(Bm, mul(Am, mul(Om, mul(Am, Bm))))
        m11, m12, m21, m22 = M
        BA = mul(Bm, Am)
        b1, b2 = BA[1] * c2, BA[3] * c2
    # C = M C M^T + b b^T for symmetric C = [[x, y], [y, z]]: a 3 x 3 linear system per mode
    K = np.zeros((w2.size, 3, 3))
    K[:, 0, 0] = 1 - m11 * m11
    K[:, 0, 1] = -2 * m11 * m12
    K[:, 0, 2] = -m12 * m12
    K[:, 1, 0] = -m11 * m21
    K[:, 1, 1] = 1 - (m11 * m22 + m12 * m21)
    K[:, 1, 2] = -m12 * m22
    K[:, 2, 0] = -m21 * m21
    K[:, 2, 1] = -2 * m21 * m22
    K[:, 2, 2] = 1 - m22 * m22
    rhs = np.stack([b1 * b1, b1 * b2, b2 * b2], axis=1)
    sol = np.linalg.solve(K, rhs[..., None])[..., 0]
    return float(sol[:, 2].mean())


def linear_table(c):
    x = np.linspace(0.0, 1e7, 65)
    return c * x, 0.0, 1.0 / (x[1] - x[0])


def run_to_equilibrium(scheme, T, samples=16, burn=1000, every=200):
    from tip.lattice import Lattice
    from tip import potential as P
    dt, dx, eta = MU * DT_PHYS, MU * DX_PHYS, ETA_PHYS / MU
    shape = (N, N, N)
    lat = Lattice(shape, np.float32)
    acc, spec = 0.0, None
    try:
        lat.upload("phi", np.zeros(shape, np.float32))
        lat.upload("pi", np.zeros(shape, np.float32))
        lat.set_table(*linear_table(M2 * MU ** 2), odd=True)
        if scheme == "rk2":
            ns = P.noise_scale(eta, T, dt, MU, DX_PHYS)
            step = lambda k, s0: lat.step(k, dt=dt, dx=dx, eta_eff=eta, mu=MU, noise_scale=ns,
                                          seed=7, step0=s0)
            field = "pi"
        elif scheme == "baoab":
            ns = math.sqrt(T / (MU ** 2 * DX_PHYS ** 3))
            step = lambda k, s0: lat.step_baoab(k, dt=dt, dx=dx, gamma=eta, mu=MU,
                                                noise_scale=ns, seed=7, step0=s0)
            field = "pi"
        else:
            ns = math.sqrt(2.0 * T * dt / (eta * MU ** 2 * DX_PHYS ** 3))
            step = lambda k, s0: lat.step_overdamped(k, dt=dt, dx=dx, eta_eff=eta, mu=MU,
                                                     noise_scale=ns, seed=7, step0=s0)
            field = "phi"
        n = 0
        step(burn, n)
        n += burn
        for _ in range(samples):
            step(every, n)
            n += every
            f = lat.download(field).astype(np.float64)
            acc += float(np.mean(f * f))
            if field == "pi":
                pk = np.abs(np.fft.fftn(f)) ** 2
                spec = pk if spec is None else spec + pk
    finally:
        lat.close()
    return acc / samples, ns, spec


@pytest.mark.parametrize("scheme", ["rk2", "baoab"])
def test_kinetic_equipartition(scheme):
    dt, dx, eta = MU * DT_PHYS, MU * DX_PHYS, ETA_PHYS / MU
    lam = lam_grid(N, dx)
    ratios = []
    for T in (7350.0, 3675.0):
        target = T / (MU ** 2 * DX_PHYS ** 3)
        pi2, ns, spec = run_to_equilibrium(scheme, T)
        exact = stationary_pi_var(scheme, lam, dt, eta, ns)
        print(f"[{scheme} T={T}] <pi^2>/target = {pi2 / target:.4f}, "
              f"exact scheme value / target = {exact / target:.4f}", flush=True)
        assert abs(pi2 / target - 1.0) < 0.10          # reference bar
        assert abs(pi2 / exact - 1.0) < 0.01            # the scheme's exact stationary value
        ratios.append(pi2 / target)
        if T == 7350.0:
            # white velocity spectrum: shell averages of |pi_k|^2 over the sub-Nyquist band
            kx = np.fft.fftfreq(N) * 2 * np.pi / dx
            kk = np.sqrt(kx[:, None, None] ** 2 + kx[None, :, None] ** 2
                         + kx[None, None, :] ** 2)
            k_ir, k_nyq = 2 * np.pi / (N * dx), np.pi / dx
            edges = np.linspace(4 * k_ir, 0.95 * k_nyq, 16)
            idx = np.digitize(kk.ravel(), edges)
            shells = [spec.ravel()[idx == b].mean() for b in range(1, len(edges))]
            spread = float(np.std(shells) / np.mean(shells))
            print(f"[{scheme}] |pi_k|^2 shell spread = {spread:.4f}", flush=True)
            assert spread < 0.25
    assert abs(ratios[0] - ratios[1]) < 0.12


def test_overdamped_equipartition():
    """Overdamped: <phi^2> = mode average of T / (mu^2 dx^3 (lam_k + M^2)) in the continuum;
    the Euler-Maruyama scheme's exact stationary value carries 1 / (1 - dt G_k / 2),
    G_k = (lam_k + M^2) / eta."""
    dt, dx, eta = MU * DT_PHYS, MU * DX_PHYS, ETA_PHYS / MU
    lam = lam_grid(N, dx)
    for T in (7350.0, 3675.0):
        base = T / (MU ** 2 * DX_PHYS ** 3)
        cont = float(np.mean(base / (lam + M2)))
        G = (lam + M2) / eta
        exact = float(np.mean(base / (lam + M2) / (1.0 - 0.5 * dt * G)))
        phi2, _, _ = run_to_equilibrium("overdamped", T, burn=1500)
        print(f"[overdamped T={T}] <phi^2>/continuum = {phi2 / cont:.4f}, "
              f"/exact = {phi2 / exact:.4f}", flush=True)
        assert abs(phi2 / cont - 1.0) < 0.10
        assert abs(phi2 / exact - 1.0) < 0.01
