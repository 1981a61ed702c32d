"""The shipped surrogate weights (tip/data/surrogate_cooling_6500_7500.npz, trained on the GPU by
tip.surrogate.train_cooling_surrogate) evaluated by the float64 oracle network on the CPU: V' against the
float64 potential within the floored 1e-4 bar (tools/validate_langevin.py:137-141 denominator)
on fresh samples, and the stored provenance."""
import json

import numpy as np

from oracle import mlp_oracle as M
from tip import surrogate as S


def test_shipped_weights_meet_the_bar():
    spec, params = S.load_params_npz(S.TRAINED_COOLING)
    assert params.shape == (S.NPARAMS,) and params.dtype == np.float32
    rho, T = spec.sample(1 << 15, seed=1234)
    y, _ = M.forward(params.astype(np.float64), spec.x(rho, T))
    got = spec.vprime_from_net(rho, T, y)
    truth = spec.vprime_truth(rho, T)
    rel = np.abs(got - truth) / np.maximum(np.abs(truth), spec.floor(T))
    print(f"max floored rel error {rel.max():.3e}")
    assert rel.max() <= 1e-4
    meta = json.loads(str(np.load(S.TRAINED_COOLING, allow_pickle=False)["meta"]))
    assert meta["recipe"].startswith("tip.surrogate.train_cooling_surrogate")


def test_shipped_weights_effective_mass():
    """The effective mass M_eff^2 = V'/rho at rho -> 0 -- what decides the spinodal instability
    near T_c -- within 1 % of the float64 potential at every T of the cooling range up to 7350 K
    (where M_eff^2 is only 1.4e-3 m^2), and the V'/rho error in the barrier region
    (rho < 2e4 GeV) within 5e-4 m^2."""
    spec, params = S.load_params_npz(S.TRAINED_COOLING)
    m2 = spec.model.mphi ** 2
    for Tv in (6500.0, 6770.0, 7000.0, 7200.0, 7300.0, 7350.0):
        r0 = np.array([spec.RHO_ZERO])
        y, _ = M.forward(params.astype(np.float64), spec.x(r0, np.array([Tv])))
        got = spec.vprime_from_net(r0, np.array([Tv]), y)[0] / spec.RHO_ZERO
        want = spec.meff2(np.array([Tv]))[0]
        assert abs(got - want) <= 1e-2 * abs(want), (Tv, got / m2, want / m2)
        r = np.logspace(0, np.log10(2e4), 300)
        Tr = np.full_like(r, Tv)
        yb, _ = M.forward(params.astype(np.float64), spec.x(r, Tr))
        d = np.abs(spec.vprime_from_net(r, Tr, yb) - spec.vprime_truth(r, Tr)) / r / m2
        assert d.max() <= 5e-4, (Tv, d.max())
