"""The learned V' surrogate under continuous cooling against the table method fed the same
float64 potential (-m gpu): the reduced form of the cooling-run acceptance.

Workload: complex field 64^3, RK2, linear cooling 7350 -> 6770 K compressed into 1000 steps
(the transition -- M_eff^2 turning negative near 7345 K, the roll to the true vacuum, the
string network decaying -- happens inside the run), 4 seeds per method, reports every 100
steps.  "table_exact": the reference's table method (16384 points on the field range, rebuilt
when |dT|/T > 1e-4 or the field leaves the range, latticeSimeRescale_numba.py:3650-3672) on
the float64 potential the network was trained on (series J + Coleman-Weinberg).  "surrogate":
the shipped network inline at the exact T(t), no tables.

Bar (the surrogate's acceptance): the true-vacuum fraction and the string count agree within the ensembles'
combined 1 sigma at every report step; the run covers the transition (f_true goes from 0 to 1
and the string count falls by more than 100x).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_surrogate_matches_exact_table_under_cooling():
    from tip import experiments as E
    from tip.surrogate import Surrogate
    res = E.surrogate_vs_table(Surrogate.load(), n=64, steps=1000, seeds=(0, 1, 2, 3),
                               report=100, cool_steps=1000, methods=("table_exact", "surrogate"))
    te, su = res["table_exact"], res["surrogate"]
    for name, d in (("table_exact", te), ("surrogate", su)):
        print(f"{name}: f_true {np.round(d['f_mean'], 4).tolist()} strings "
              f"{np.round(d['s_mean']).astype(int).tolist()} ({d['ms_per_step']:.3f} ms/step, "
              f"{d['table_rebuilds']} table rebuilds)", flush=True)
    assert su["table_rebuilds"] == 0 and te["table_rebuilds"] > 0
    for key in ("f", "s"):
        m1, s1, m2, s2 = te[key + "_mean"], te[key + "_sd"], su[key + "_mean"], su[key + "_sd"]
        tol = np.maximum(np.sqrt(s1 ** 2 + s2 ** 2), 1e-12)
        assert np.all(np.abs(m1 - m2) <= tol), (key, m1, m2, tol)
    assert te["f_mean"][0] == 0.0 and te["f_mean"][-1] == 1.0
    assert te["s_mean"][0] > 100 * te["s_mean"][-1]
