"""Pins of the CPU surrogate-network oracle and of the surrogate's host logic (CPU).

  * oracle gradients against central finite differences (a wrong transpose or a dropped
    (1 - h^2) factor fails)
  * the float64 ground truth: V'/rho - tree is the thermal + CW part; at rho -> 0 it tends to
    the thermal mass^2 / m^2 given by the high-temperature limit of the dJ functions
  * feature map range and the sampling mixture
"""
import math

import numpy as np
import pytest

from oracle import mlp_oracle as M


def test_oracle_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    p = M.random_params(3).astype(np.float64)
    X = rng.uniform(-1, 1, (37, 2))
    Y = rng.normal(0, 0.5, 37)
    W = rng.uniform(0.1, 3.0, 37)
    L, g = M.loss_grad(p, X, Y, W)
    idx = rng.choice(M.NP, 60, replace=False)
    for i in list(idx) + [0, 64, 96, 1120, 1152, 1184]:
        e = np.zeros(M.NP)
        h = 1e-6
        e[i] = h
        fd = (M.loss_grad(p + e, X, Y, W)[0] - M.loss_grad(p - e, X, Y, W)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-7 + 1e-6 * abs(g[i]), i


def test_oracle_forward_shapes_and_layout():
    p = np.zeros(M.NP)
    p[-1] = 0.25  # b3 only
    y, _ = M.forward(p, np.zeros((5, 2)))
    assert np.allclose(y, 0.25)
    W1, b1, W2, b2, w3, b3 = M.unpack(np.arange(M.NP, dtype=float))
    assert W1[0, 1] == 1 and b1[0] == 64 and W2[0, 0] == 96 and b2[0] == 1120
    assert w3[0] == 1152 and b3 == 1184


def test_surrogate_target_is_thermal_part():
    from tip import potential as P
    from tip import surrogate as S
    spec = S.SurrogateSpec()
    rho = np.array([1e-1, 1e2, 1e4, 1e6, 1e9, 1e11])
    T = np.full_like(rho, 7000.0)
    g = spec.g_full(rho, T)
    # Boltzmann suppression: nothing thermal left far above T
    assert abs(g[3]) < 1e-10 and abs(g[4]) < 1e-10
    p = spec.model

    def lim(Tv):
        """rho -> 0: g -> (1/m^2) T^2/(2 pi^2) sum_s c_s dJ_s(u_s)/u_s * y_s^2/2 (the chain-rule
        limit of the thermal part; the Coleman-Weinberg part vanishes as rho^2 log rho)."""
        ub = math.sqrt(p.mb2 + (0.25 * p.yb ** 2 + 2 / 3 * p.gb ** 2) * Tv ** 2) / Tv
        uf = math.sqrt((1 / 6) * p.gf ** 2 * Tv ** 2) / Tv
        return Tv ** 2 / (2 * math.pi ** 2) * (
            p.coeff_b * P.dJb(ub)[0] / ub * 0.5 * p.yb ** 2
            + p.coeff_f * P.dJf(uf)[0] / uf * 0.5 * p.yf ** 2) / p.mphi ** 2

    assert abs(g[0] - lim(7000.0)) < 1e-6 * abs(lim(7000.0))
    # the effective-mass fit g0(T) the kernels use is that limit across the range
    for Tv in (6500.0, 6770.0, 7000.0, 7350.0, 7500.0):
        assert abs(spec.g0_of(Tv) - lim(Tv)) < 1e-7, Tv
    # the network's target is the rest of g over f1
    y = spec.g_target(rho, T)
    f1 = spec.x(rho, T)[:, 0]
    np.testing.assert_allclose(spec.g0_of(T) + f1 * y, g, rtol=1e-12, atol=1e-12)


def test_features_and_sampling():
    from tip import surrogate as S
    spec = S.SurrogateSpec()
    rho, T = spec.sample(4096, seed=1)
    X = spec.x(rho, T)
    assert X.shape == (4096, 2)
    assert X[:, 0].min() >= 0 and X[:, 0].max() <= 1.0 + 1e-12
    assert X[:, 1].min() >= -1 and X[:, 1].max() <= 1
    # the mixture puts a sizeable share of samples inside the thermal region
    assert np.mean(rho < 1e5) > 0.15
    assert np.mean(rho > 0.1 * spec.rho_max) > 0.4
