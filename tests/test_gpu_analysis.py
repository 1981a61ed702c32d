"""Bubble and string component analysis on the GPU (-m gpu) against the oracle
(oracle/ccl_oracle.py: the reference's _identify_bubbles / _identify_strings on
scipy.ndimage.label, postprocess/revisualize_snapshots.py:171-213, 360-392).

Bars: labels and component counts bit-identical to scipy's (numbering included); volumes,
centroids, r_eff, peaks, signs and order identical; r_rms within 1e-12 relative; the periodic
variant equal to the breadth-first oracle with wrap-around; full size 256^3 labels identical.
"""
import numpy as np
import pytest
from scipy import ndimage

from oracle import ccl_oracle as CO
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def smooth_field(shape, seed, sigma=2.0, dtype=np.float32):
    rng = np.random.default_rng(seed)
    f = ndimage.gaussian_filter(rng.normal(size=shape), sigma, mode="wrap")
    return (f / f.std() * 1e4).astype(dtype)


@pytest.mark.parametrize("shape", [(17, 23, 70), (64, 64, 64), (1, 1, 5), (3, 1, 130)])
@pytest.mark.parametrize("p", [0.1, 0.3116, 0.6])
def test_labels_equal_scipy(shape, p):
    from tip import analysis as A
    rng = np.random.default_rng(int(p * 1e4) + shape[2])
    v = rng.random(shape).astype(np.float32)
    lab, n = A.label(v, 1.0 - p)
    ref, nref = CO.label(v > np.float32(1.0 - p))
    assert n == nref and np.array_equal(lab, ref)


def test_labels_periodic_equal_bfs():
    from tip import analysis as A
    rng = np.random.default_rng(3)
    v = rng.random((9, 10, 11))
    lab, n = A.label(v, 0.7, periodic=True)
    ref, nref = CO.label_bfs(v > 0.7, periodic=True)
    assert n == nref and np.array_equal(lab, ref)


def test_empty_and_full():
    from tip import analysis as A
    z = np.zeros((8, 9, 10), np.float32)
    lab, n = A.label(z, 0.5)
    assert n == 0 and not lab.any()
    assert A.identify_bubbles(z, 0.5) == []
    lab, n = A.label(z + 1.0, 0.5)
    assert n == 1 and (lab == 1).all()


def _same_components(got, ref, keys_exact, key_close=None):
    assert len(got) == len(ref)
    for g, r in zip(got, ref):
        for k in keys_exact:
            if isinstance(r[k], np.ndarray):
                assert np.array_equal(g[k], r[k]), k
            else:
                assert g[k] == r[k], k
        if key_close:
            assert abs(g[key_close] - r[key_close]) <= 1e-12 * max(abs(r[key_close]), 1.0)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_bubbles_match_oracle(dtype):
    from tip import analysis as A
    phi = smooth_field((48, 40, 72), 5, dtype=dtype)
    got = A.identify_bubbles(phi, 1.2e4, dx_phys=0.002)
    ref = CO.identify_bubbles(phi, 1.2e4, dx_phys=0.002)
    assert len(ref) > 20
    _same_components(got, ref, ["volume", "centroid", "r_eff", "sign", "peak", "r_phys"],
                     "r_rms")
    s = A.bubble_summary(got, phi.size)
    assert s["n_bub"] == len(ref) and s["tot_vol"] == sum(b["volume"] for b in ref)


def test_strings_match_oracle():
    from tip import analysis as A
    from tip.lattice import Lattice
    shape = (40, 48, 64)
    p1, p2 = smooth_field(shape, 1, 1.5), smooth_field(shape, 2, 1.5)
    w_ref, _ = O.winding(p1, p2)
    w = np.rint(w_ref).astype(np.float32)
    lab, got = A.identify_strings(w)
    lref, ref = CO.identify_strings(w)
    assert len(ref) > 10 and np.array_equal(lab, lref)
    _same_components(got, ref, ["loop_id", "n_voxels", "centroid", "winding_sign",
                                "max_winding"])
    lat = Lattice(shape, np.float32, complex_field=True)
    try:
        lat.upload("phi1", p1)
        lat.upload("phi2", p2)
        lab2, got2 = lat.strings()
    finally:
        lat.close()
    assert np.array_equal(lab2, lref)
    _same_components(got2, ref, ["loop_id", "n_voxels", "centroid", "winding_sign"])


def test_lattice_bubbles_full_size():
    """256^3 real field near the percolation threshold: labels identical to scipy's."""
    from tip import analysis as A
    from tip.lattice import Lattice
    shape = (256, 256, 256)
    phi = smooth_field(shape, 9, 1.5)
    lab, n = A.label(phi, 5e3)
    ref, nref = CO.label((phi > 5e3) | (phi < -5e3))
    assert n == nref and np.array_equal(lab, ref)
    lat = Lattice(shape, np.float32)
    try:
        lat.upload("phi", phi)
        got = lat.bubbles(5e3)
    finally:
        lat.close()
    assert len(got) == nref and got[0]["volume"] == int(np.bincount(ref.ravel())[1:].max())
