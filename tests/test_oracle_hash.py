"""Pins of the oracle's hash RNG (CPU).

The reference's hash lives in two independent texts: the Numba kernel
(simulation/latticeSimeRescale_numba.py:2322-2336) and the C++ evolver
(cosmolattice_ext/evolvers/stochasticrk.h:206-225).  The oracle's C restatement is checked
against (a) a pure-Python big-integer restatement, (b) a numpy uint64 restatement, (c) hand
values of the splitmix finaliser, and (d) the statistical properties the Box-Muller output
must have.  A dropped xorshift, a wrong constant or a wrong shift breaks (a)-(c).
"""
import math

import numpy as np
import pytest

from oracle import oracle as O

EDGE_SEEDS = [0, 1, 2, 3, 73856093, 0x9E3779B97F4A7C15, (1 << 63), (1 << 64) - 1,
              (1 << 32) - 1, 1 << 32, 123456789012345678]


def test_c_matches_pure_python_on_edge_and_random_seeds():
    rng = np.random.default_rng(7)
    seeds = np.array(EDGE_SEEDS + [int(v) for v in rng.integers(0, 2**63, 500, dtype=np.int64)],
                     dtype=np.uint64)
    x1, x2, u1, u2 = O.hash_pairs(seeds)
    for i, s in enumerate(seeds.tolist()):
        px1, px2, pu1, pu2 = O.hash_pair_py(int(s))
        assert int(x1[i]) == px1 and int(x2[i]) == px2, s
        assert u1[i] == pu1 and u2[i] == pu2, s


def test_c_matches_numpy_restatement_bitwise():
    seeds = O.site_seeds(1 << 16, noise_seed=77)
    a = O.hash_pairs(seeds)
    b = O.hash_pairs_np(seeds)
    for u, v in zip(a, b):
        assert np.array_equal(u, v)


def test_first_mixer_is_splitmix64_finaliser():
    # The first mixer is the splitmix64 finaliser; splitmix64's first output for state 0 is
    # mix(0 + 0x9E3779B97F4A7C15) = 0xE220A8397B1DCDAF (published sequence of seed 0).
    x1, _, _, _ = O.hash_pair_py(0x9E3779B97F4A7C15)
    assert x1 == 0xE220A8397B1DCDAF


def test_u_range_and_clamp():
    seeds = O.site_seeds(1 << 18, noise_seed=5)
    _, _, u1, u2 = O.hash_pairs(seeds)
    assert u1.min() >= 1e-12 and u1.max() < 1.0
    assert u2.min() >= 0.0 and u2.max() < 1.0
    # seed 0 mixes to 0 -> u1 = 0 -> clamped to 1e-12 (the reference's guard)
    _, _, a, b = O.hash_pair_py(0)
    assert a == 1e-12 and b > 0.0


def test_site_seed_formula():
    # uint64(site*73856093) ^ noise_seed (latticeSimeRescale_numba.py:2514-2516)
    s = O.site_seeds(10, noise_seed=3)
    assert [int(v) for v in s] == [(i * 73856093) ^ 3 for i in range(10)]


def test_gaussian_moments_and_ks():
    from scipy import stats

    z = O.noise_field((32, 32, 64), noise_seed=11, scale=1.0)
    z = z.ravel()
    n = z.size
    assert abs(z.mean()) < 5.0 / math.sqrt(n)
    assert abs(z.var() - 1.0) < 5.0 * math.sqrt(2.0 / n)
    assert stats.kstest(z, "norm").pvalue > 1e-4


def test_pair_components_are_cos_sin_of_same_draw():
    z1, z2 = O.noise_field((4, 8, 16), noise_seed=9, scale=2.0, pair=True)
    seeds = O.site_seeds(4 * 8 * 16, 9)
    _, _, u1, u2 = O.hash_pairs_np(seeds)
    r = np.sqrt(-2.0 * np.log(u1))
    np.testing.assert_allclose(z1.ravel(), 2.0 * r * np.cos(2 * np.pi * u2), rtol=1e-14, atol=1e-14)
    np.testing.assert_allclose(z2.ravel(), 2.0 * r * np.sin(2 * np.pi * u2), rtol=1e-14, atol=1e-14)
    # the two components are uncorrelated (5 sigma on 65536 pairs)
    a, b = O.noise_field((32, 32, 64), noise_seed=10, scale=1.0, pair=True)
    assert abs(np.corrcoef(a.ravel(), b.ravel())[0, 1]) < 5.0 / math.sqrt(a.size)


def test_noise_seed_fold():
    assert O.noise_seed(123, 0) == 123
    assert O.noise_seed(123, 1) == 123 ^ 0x9E3779B97F4A7C15


REF_HASH = O._LIB_PATH.parent / "_ref" / "ref_hash_probe"
_M = (1 << 64) - 1


def _unmix(x1: int) -> int:
    """The seed whose first-mixer state is x1 (the splitmix64 finaliser is a bijection)."""
    def unxs(y, k):
        x = y
        for _ in range(64 // k + 1):
            x = y ^ (x >> k)
        return x
    x = unxs(x1, 31)
    x = (x * pow(0x94D049BB133111EB, -1, 1 << 64)) & _M
    x = unxs(x, 27)
    x = (x * pow(0xBF58476D1CE4E5B9, -1, 1 << 64)) & _M
    return unxs(x, 30)


@pytest.mark.skipif(not REF_HASH.exists(), reason="oracle/_ref not built (no reference checkout)")
def test_u1_u2_equal_reference_cpp_bitwise():
    """u1 and u2 bit for bit against the reference's own C++ text of the hash
    (hashMix64 / hashRngPair, cosmolattice_ext/evolvers/stochasticrk.h:205-225, compiled from
    the checkout by oracle/build_ref.py): edge seeds, 2^16 site seeds of two steps, and seeds
    whose u1 falls under the 1e-12 clamp.  This pins the second mixer (the source of u2),
    which the splitmix64 identity does not cover."""
    import subprocess
    seeds = [int(s) for s in EDGE_SEEDS]
    for step in (0, 12345):
        seeds += [int(v) for v in O.site_seeds(1 << 16, noise_seed=O.noise_seed(step, 3))]
    # seeds whose u1 falls under the clamp: the first mixer inverted at small states x1
    seeds += [_unmix(x1) for x1 in (0, 1, 12345 << 11, (9007 << 11) - 1)]
    out = subprocess.run([str(REF_HASH)], input="\n".join(map(str, seeds)) + "\n",
                         capture_output=True, text=True, check=True).stdout.split()
    ref = np.array([int(v, 16) for v in out], dtype=np.uint64).reshape(-1, 2)
    _, _, u1, u2 = O.hash_pairs(np.array(seeds, dtype=np.uint64))
    assert ref.shape == (len(seeds), 2)
    assert np.array_equal(u1.view(np.uint64), ref[:, 0])
    assert np.array_equal(u2.view(np.uint64), ref[:, 1])
    assert np.sum(u1 == 1e-12) >= 3  # the clamp cases are in the sample
