"""Pins of the component-analysis oracle (CPU): scipy.ndimage.label (the routine the reference
calls) against an independent breadth-first labeller -- 6-connectivity (face neighbours join,
edge and corner neighbours do not), no wrap-around, numbering in raster order of each
component's first site -- and the per-component measures against hand-computed values."""
import numpy as np
import pytest

from oracle import ccl_oracle as CO


@pytest.mark.parametrize("p", [0.15, 0.3116, 0.5])
def test_scipy_label_equals_bfs(p):
    rng = np.random.default_rng(int(p * 1000))
    m = rng.random((7, 9, 11)) < p
    a, na = CO.label(m)
    b, nb = CO.label_bfs(m)
    assert na == nb and np.array_equal(a, b)


def test_connectivity_and_boundary():
    m = np.zeros((4, 4, 4), bool)
    m[0, 0, 0] = m[1, 1, 0] = True          # edge neighbours: two components
    m[3, 3, 3] = m[3, 3, 2] = True          # face neighbours: one component
    m[0, 3, 3] = True                       # touches the far corner only through the wrap
    lab, n = CO.label(m)
    assert n == 4
    assert lab[0, 0, 0] == 1 and lab[0, 3, 3] == 2 and lab[1, 1, 0] == 3
    assert lab[3, 3, 2] == lab[3, 3, 3] == 4
    # periodic: [0, 3, 3] and [3, 3, 3] are x-neighbours through the wrap
    lp, npp = CO.label_bfs(m, periodic=True)
    assert npp == 3 and lp[0, 3, 3] == lp[3, 3, 3]


def test_bubble_measures_hand_values():
    phi = np.zeros((6, 6, 6), np.float32)
    phi[1:3, 1:3, 1:3] = 2e4            # 2x2x2 cube, centroid (1.5, 1.5, 1.5)
    phi[4, 4, 4] = -3e4                 # single site
    phi[5, 0, 0] = -2e4
    phi[5, 0, 1] = 1e4 + 1              # row of 2, mean > 0
    b = CO.identify_bubbles(phi, 1e4)
    assert [x["volume"] for x in b] == [8, 2, 1]
    assert np.allclose(b[0]["centroid"], 1.5) and b[0]["r_rms"] == pytest.approx(np.sqrt(0.75))
    assert b[0]["r_eff"] == pytest.approx((3 * 8 / (4 * np.pi)) ** (1 / 3))
    assert b[1]["sign"] == -1 and b[1]["peak"] == 2e4
    assert b[2]["sign"] == -1 and b[2]["peak"] == 3e4 and b[2]["r_rms"] == 0.0


def test_string_measures_hand_values():
    w = np.zeros((5, 5, 5), np.float32)
    w[1, 1, :] = 1.0                    # a line along z: 5 voxels
    w[3, 3, 3] = -2.0
    lab, s = CO.identify_strings(w)
    assert [x["n_voxels"] for x in s] == [5, 1]
    assert s[0]["winding_sign"] == 1 and s[1]["winding_sign"] == -1
    assert s[1]["max_winding"] == 2.0 and np.allclose(s[0]["centroid"], [1, 1, 2])
