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
