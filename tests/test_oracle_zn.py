"""Pins of the oracle's Z_N force (CPU).

The complex step adds dV1 += -N dV sin(N theta) phi2 / rho^2, dV2 += +N dV sin(N theta) phi1 /
rho^2 (simulation/latticeSimComplex_numba.py:2147-2152; oracle/tip_oracle.c complex_force).
Three independent checks, each failing for a flipped sign, a swapped component or a wrong
factor of N:
  * the term is the gradient of V_ZN = -dV cos(N atan2(phi2, phi1)) (central differences);
  * it equals the reference's own C++ statement of the same force,
    ThermalTables::vPrimeComponents (cosmolattice_ext/models/thermal_tables.hpp:161-178),
    compiled from the checkout into oracle/_ref by oracle/build_ref.py;
  * one full oracle step with Z_N on commutes with a rotation of (phi, pi, noise) by 2 pi / N
    (the residual Z_N symmetry), and not with a generic rotation.
"""
import math
import subprocess
from pathlib import Path

import numpy as np
import pytest

from oracle import oracle as O

ROOT = Path(__file__).resolve().parents[1]
PROBE = ROOT / "oracle" / "_ref" / "ref_zn_probe"
ZERO = np.zeros(16)  # a zero V'(rho) table: the force is the Z_N term alone


def _points(rng, n=400):
    r = 10.0 ** rng.uniform(1.0, 5.0, n)
    th = rng.uniform(-math.pi, math.pi, n)
    p1, p2 = r * np.cos(th), r * np.sin(th)
    # the negative real axis (theta = +-pi, signed zero), the axes, tiny radii
    extra = np.array([[-500.0, 0.0], [-500.0, -0.0], [0.0, 300.0], [0.0, -300.0], [700.0, 0.0],
                      [1e-19, 2e-19], [-3e-5, 4e-5]])
    return np.concatenate([p1, extra[:, 0]]), np.concatenate([p2, extra[:, 1]])


@pytest.mark.parametrize("n", [3.0, 4.0, 6.0, 2.5])
def test_zn_force_is_gradient_of_cos_potential(rng, n):
    dv = 2.5e9
    p1, p2 = _points(rng)
    keep = (np.hypot(p1, p2) > 1.0) & (np.abs(p2) > 1e-6 * np.abs(p1))  # away from the cut
    p1, p2 = p1[keep], p2[keep]
    f1, f2 = O.complex_force(p1, p2, ZERO, 0.0, 1.0, zn=(n, dv, 1.0))

    def V(a, b):
        return -dv * np.cos(n * np.arctan2(b, a))

    h = 1e-5 * np.hypot(p1, p2)
    g1 = (V(p1 + h, p2) - V(p1 - h, p2)) / (2 * h)
    g2 = (V(p1, p2 + h) - V(p1, p2 - h)) / (2 * h)
    scale = n * dv / np.hypot(p1, p2)
    assert np.max(np.abs(f1 - g1) / scale) < 1e-7
    assert np.max(np.abs(f2 - g2) / scale) < 1e-7
    # the term switches off where the reference switches it off
    a, b = O.complex_force(np.array([1e-21, 3.0]), np.array([0.0, 4.0]), ZERO, 0.0, 1.0,
                           zn=(n, dv, 0.0))
    assert not np.any(a) and not np.any(b)


@pytest.mark.skipif(not PROBE.exists(), reason="oracle/_ref not built (no reference checkout)")
def test_zn_force_equals_reference_cpp(rng, tmp_path):
    from test_potential import _write_bin
    binp = tmp_path / "thermal_tables.bin"
    _write_bin(binp)
    p1, p2 = _points(rng)
    rows = []
    for n in (3.0, 4.0, 6.0):
        for T in (6770.0, 7350.0):
            dv = 1e13 / n
            for a, b in zip(p1, p2):
                rows.append((a, b, T, n, dv))
    inp = "\n".join(" ".join(repr(float(v)) for v in r) for r in rows) + "\n"
    out = subprocess.run([str(PROBE), str(binp)], input=inp, capture_output=True, text=True,
                         check=True).stdout.split("\n")
    vals = np.array([[float(v) for v in line.split()] for line in out if line.strip()])
    assert vals.shape == (len(rows), 4)
    rows = np.array(rows)
    checked = 0
    for n in (3.0, 4.0, 6.0):
        sel = rows[:, 3] == n
        f1, f2 = O.complex_force(rows[sel, 0], rows[sel, 1], ZERO, 0.0, 1.0,
                                 zn=(n, 1e13 / n, 1.0))
        on, off = vals[sel, :2], vals[sel, 2:]
        ref1, ref2 = on[:, 0] - off[:, 0], on[:, 1] - off[:, 1]
        tol = 1e-12 * (np.abs(on) + np.abs(off)).max(axis=1) + 1e-300
        assert np.all(np.abs(f1 - ref1) <= tol) and np.all(np.abs(f2 - ref2) <= tol), n
        checked += int(sel.sum())
    assert checked == len(rows)


def _step(fields, noise, t, zn):
    shape = fields[0].shape
    got = [f.copy() for f in fields]
    tmp = [np.empty(shape) for _ in range(4)]
    O.rk2_step_table_complex(*got, 0.25, 1.0, 1.0, 1000.0, noise[0], noise[1], *tmp, 1.0, t, 0.0,
                             20 / 4000.0, len(t), zn=zn)
    return got


@pytest.mark.parametrize("n", [3, 4])
def test_zn_step_commutes_with_residual_rotation(rng, n):
    shape = (4, 4, 6)
    f = [rng.normal(size=shape) * 900.0 for _ in range(2)]
    p = [rng.normal(size=shape) * 200.0 for _ in range(2)]
    nz = [rng.normal(size=shape) * 40.0 for _ in range(2)]
    t = np.abs(rng.normal(size=21)) * 1e5
    zn = (float(n), 4e10, 1.0)

    def rot(a, x, y):
        return math.cos(a) * x - math.sin(a) * y, math.sin(a) * x + math.cos(a) * y

    def check(a):
        r1 = _step([*rot(a, *f), *rot(a, *p)], rot(a, *nz), t, zn)
        r0 = _step([*f, *p], nz, t, zn)
        e = [*rot(a, r0[0], r0[1]), *rot(a, r0[2], r0[3])]
        return max(float(np.max(np.abs(g - x) / (np.abs(x).max()))) for g, x in zip(r1, e))

    assert check(2 * math.pi / n) < 1e-12
    assert check(2 * math.pi * 3 / n) < 1e-12
    assert check(0.7) > 1e-6  # the Z_N term breaks the full U(1)
