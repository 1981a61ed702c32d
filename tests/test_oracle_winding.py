"""Pins of the oracle's plaquette winding and vacuum count (CPU).

  * an analytic vortex line phi = (x - x0) + i (y - y0) pierces exactly the XY plaquettes
    around (x0, y0) with winding +1; XZ/YZ plaquettes carry none
  * a uniform phase gives zero everywhere
  * divergence invariant: the outward-oriented sum over the six faces of every unit cube
    vanishes (XY normal +z, XZ normal -y, YZ normal +x) -- a wrong corner order or a wrong
    orientation of one plaquette family breaks it
  * an independent numpy np.roll restatement (the structure of tools/winding.py:26-64)
  * vacuum count against a brute-force Python count
"""
import numpy as np
import pytest

from oracle import oracle as O


def _winding_np(p1, p2):
    th = np.arctan2(p2, p1)
    wrap = lambda d: np.arctan2(np.sin(d), np.cos(d))
    pw = lambda a, b, c, d: (wrap(b - a) + wrap(c - b) + wrap(d - c) + wrap(a - d)) / (2 * np.pi)
    ip = np.roll(th, -1, 0)
    jp = np.roll(th, -1, 1)
    kp = np.roll(th, -1, 2)
    return (pw(th, ip, np.roll(ip, -1, 1), jp), pw(th, ip, np.roll(ip, -1, 2), kp),
            pw(th, jp, np.roll(jp, -1, 2), kp))


def test_vortex_line():
    nx, ny, nz = 8, 6, 5
    x0, y0 = 3.5, 2.5
    i, j, k = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    p1 = (i - x0).astype(np.float64)
    p2 = (j - y0).astype(np.float64)
    w, plaq = O.winding(p1, p2)
    interior = plaq[0, : nx - 1, : ny - 1, :]
    expect = np.zeros_like(interior)
    expect[3, 2, :] = 1
    assert np.array_equal(interior, expect)
    assert np.all(plaq[1] == 0) and np.all(plaq[2] == 0)


def test_uniform_phase_zero():
    p1 = np.full((4, 5, 6), 2.0)
    p2 = np.full((4, 5, 6), -1.0)
    w, plaq = O.winding(p1, p2)
    assert np.all(plaq == 0) and np.all(w == 0)
    assert O.string_count(w) == 0


def test_cube_divergence_free(rng):
    shape = (7, 6, 9)
    p1 = rng.normal(size=shape).astype(np.float32)
    p2 = rng.normal(size=shape).astype(np.float32)
    _, plaq = O.winding(p1, p2)
    wxy, wxz, wyz = (plaq[o].astype(np.int64) for o in range(3))
    div = (np.roll(wxy, -1, 2) - wxy) - (np.roll(wxz, -1, 1) - wxz) + (np.roll(wyz, -1, 0) - wyz)
    assert np.all(div == 0)
    assert np.abs(plaq).sum() > 0  # random fields do contain strings


def test_matches_numpy_restatement(rng):
    shape = (6, 7, 8)
    p1 = rng.normal(size=shape).astype(np.float32).astype(np.float64)
    p2 = rng.normal(size=shape).astype(np.float32).astype(np.float64)
    w, plaq = O.winding(p1, p2)
    ref = _winding_np(p1, p2)
    for o in range(3):
        assert np.array_equal(plaq[o], np.rint(ref[o]).astype(np.int8))
    np.testing.assert_allclose(w, (ref[0] + ref[1] + ref[2]).astype(np.float32), atol=1e-6)
    assert O.string_count(w) == int(np.sum(np.abs(ref[0] + ref[1] + ref[2]) > 0.5))


def test_vacuum_counts(rng):
    phi = rng.normal(size=(5, 6, 7)) * 2e4
    esc = 1e4
    assert O.vacuum_count(phi, esc) == sum(1 for v in phi.ravel() if abs(v) > esc)
    f = phi.astype(np.float32)
    assert O.vacuum_count(f, esc) == sum(1 for v in f.ravel() if abs(float(v)) > esc)
    g = (rng.normal(size=(5, 6, 7)) * 2e4).astype(np.float32)
    assert O.vacuum_count(f, esc, g) == sum(
        1 for a, b in zip(f.ravel(), g.ravel()) if float(a) ** 2 + float(b) ** 2 > esc ** 2)
