"""Learned V' surrogate on the GPU (-m gpu).

Bars:
  * device forward pass vs the float64 oracle network: 2e-6 relative to the output scale
  * device gradients vs the oracle's analytic gradients (which finite differences pin):
    3e-4 of the largest component
  * GPU Adam training (the cooling-run recipe) lowers the loss > 1e4 x and meets
    |dV'|/max(|V'|, floor) <= 1e-4 against the float64 potential
  * V' assembly, on-device tabulation and the inline-MLP step agree with the oracle's use of
    the same network
  * gain over the method without it: the trained network resolves the thermal barrier that the
    8192-point table on [0, 1.5 phi_true] cannot, and follows T continuously
"""
import numpy as np
import pytest

from oracle import mlp_oracle as M
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _spec():
    from tip import surrogate as S
    return S.SurrogateSpec()


@pytest.fixture(scope="module")
def trained():
    """The cooling-run training recipe: 2^20 float64-truth samples, weighted loss, 400 epochs of
    GPU Adam (batch 1024, cosine lr 2e-3 -> 1e-6)."""
    from tip import surrogate as S
    spec = _spec()
    rho, T = spec.sample(1 << 20, seed=0, T_margin=0.05)
    sur = S.Surrogate(spec, seed=0)
    r, t = rho[:1 << 16], T[:1 << 16]
    X = spec.x(r, t).astype(np.float32)
    Y = spec.g_target(r, t).astype(np.float32)
    W = spec.weights(r, t).astype(np.float32)
    L0, _ = sur.grad(X, Y, W)
    hist = sur.train(rho, T, epochs=400, batch=1024, lr=2e-3, lr_end=1e-6)
    L1, _ = sur.grad(X, Y, W)
    return sur, L0, L1, hist


def test_forward_matches_oracle():
    from tip import surrogate as S
    rng = np.random.default_rng(1)
    p = S.init_params(5) + rng.normal(0, 0.1, S.NPARAMS).astype(np.float32)
    sur = S.Surrogate(_spec(), params=p)
    X = rng.uniform(-1, 1, (5000, 2)).astype(np.float32)
    y = sur.forward(X)
    ref, _ = M.forward(p.astype(np.float64), X.astype(np.float64))
    assert np.max(np.abs(y - ref)) <= 2e-6 * max(1.0, np.abs(ref).max())


def test_gradient_matches_oracle():
    from tip import surrogate as S
    rng = np.random.default_rng(2)
    p = S.init_params(6) + rng.normal(0, 0.1, S.NPARAMS).astype(np.float32)
    sur = S.Surrogate(_spec(), params=p)
    for n in (64, 1000, 4096):  # whole blocks, ragged last block
        X = rng.uniform(-1, 1, (n, 2)).astype(np.float32)
        Y = rng.normal(0, 1, n).astype(np.float32)
        W = rng.uniform(0.2, 2.0, n).astype(np.float32)
        L, g = sur.grad(X, Y, W)
        Lr, gr = M.loss_grad(p.astype(np.float64), X.astype(np.float64), Y.astype(np.float64),
                             W.astype(np.float64))
        assert abs(L - Lr) <= 1e-5 * abs(Lr)
        assert np.max(np.abs(g - gr)) <= 3e-4 * np.abs(gr).max(), n


@pytest.mark.timeout(600)
def test_training_meets_accuracy(trained):
    """Loss drops > 1e4 x; V' against the float64 truth within the floored 1e-4
    (tools/validate_langevin.py:137-141 denominator) on 2^17 fresh samples, and within 2e-3 of
    rho m^2 in the barrier region."""
    sur, L0, L1, hist = trained
    print(f"weighted loss on 2^16 training samples {L0:.3e} -> {L1:.3e}")
    assert L1 < L0 / 1e4
    rep = sur.error_report(n=1 << 17)
    print(rep)
    assert rep["max_rel_floored"] <= 1e-4
    assert rep["max_rel_barrier_region"] < 2e-3


def test_vprime_tabulate_and_inline_step_consistent():
    from tip import surrogate as S
    from tip.device import DeviceArray
    from tip import _lib
    spec = _spec()
    rng = np.random.default_rng(4)
    p = S.init_params(7) + rng.normal(0, 0.05, S.NPARAMS).astype(np.float32)
    sur = S.Surrogate(spec, params=p)
    rho = np.concatenate([np.geomspace(1.0, 1.4e11, 3000), rng.uniform(0, 2e4, 1000)])
    T = rng.uniform(6600, 7400, rho.size)
    vp = sur.vprime(rho, T).astype(np.float64)
    y, _ = M.forward(p.astype(np.float64), spec.x(rho, T))
    ref = spec.vprime_from_net(rho, T, y)
    scale = rho * spec.model.mphi ** 2
    assert np.max(np.abs(vp - ref) / scale) < 1e-5
    # on-device tabulation (odd, float64 out) matches pointwise evaluation
    tab = sur.tabulate(7000.0, -3e4, 3e4, 601).download()
    grid = np.linspace(-3e4, 3e4, 601)
    pt = sur.vprime(np.abs(grid), np.full(601, 7000.0)).astype(np.float64) * np.sign(grid)
    assert np.max(np.abs(tab - pt) / np.maximum(np.abs(grid) * 1e6, 1.0)) < 1e-6
    # one complex step with the inline MLP == oracle step with a fine table of the same network
    shape = (16, 24, 64)
    f1 = (rng.normal(size=shape) * 2e3).astype(np.float32)
    f2 = (rng.normal(size=shape) * 2e3).astype(np.float32)
    p1 = (rng.normal(size=shape) * 1e2).astype(np.float32)
    p2 = (rng.normal(size=shape) * 1e2).astype(np.float32)
    rmax = float(np.sqrt(f1.astype(np.float64) ** 2 + f2.astype(np.float64) ** 2).max()) * 1.2
    xs = np.linspace(0.0, rmax, 200001)
    ys, _ = M.forward(p.astype(np.float64), spec.x(xs, 7000.0))
    fine = spec.vprime_from_net(xs, 7000.0, ys)
    a = [f1.astype(np.float64), f2.astype(np.float64), p1.astype(np.float64), p2.astype(np.float64)]
    tmp = [np.empty(shape) for _ in range(4)]
    O.rk2_step_table_complex(*a, 0.25, 1.0, 1.0, 1000.0, None, None, *tmp, 1.0, fine, 0.0,
                             (len(xs) - 1) / rmax, len(xs))
    ins = [DeviceArray.from_numpy(x) for x in (f1, f2, p1, p2)]
    outs = [DeviceArray(shape, np.float32) for _ in range(4)]
    fm = spec.features()
    import ctypes as C
    _lib.call("tip_rk2_fused_c32_mlp", *(d.ptr for d in ins), *(d.ptr for d in outs), 0.25, 1.0,
              1.0, 1000.0, 1.0, sur.params.ptr, C.byref(fm), 7000.0, 0.0, 0.0, 0.0, *shape, 0, 0.0,
              None, 0.0, None)
    got = [d.download().astype(np.float64) for d in outs]
    for g, e in zip(got, a):
        assert np.linalg.norm(g - e) / np.linalg.norm(e) < 2e-6


@pytest.mark.timeout(600)
def test_gain_over_table_and_fixed_temperature(trained):
    """What the trained network buys: (1) the 8192-point table on [0, 1.5 phi_true] cannot
    resolve the thermal barrier (spacing ~1.8e7 GeV against a ~1e3 GeV barrier) -- the
    network can; (2) a table built at T_k and held while T cools drifts from the truth, the
    network evaluates at the current T."""
    from tip import potential as P
    from tip import surrogate as S
    sur = trained[0]
    spec = sur.spec
    net = sur.error_report(n=1 << 14)
    t, tmin, dinv = P.build_vprime_table(7000.0, 0.0, spec.rho_max, 8192, spec.model)
    tab = S.table_error(spec, t, tmin, dinv, 7000.0)
    print("net", net, "table", tab)
    assert net["max_rel_barrier_region"] < 0.01 * tab["max_rel_barrier_region"]
    # held table: built at 7000, evaluated at 6950 (the reference rebuilds only when
    # |dT|/T > 1e-4, a held table is what runs between rebuilds)
    r = np.geomspace(10.0, 3e4, 400)
    truth = spec.vprime_truth(r, np.full_like(r, 6950.0))
    fine_t, ft_min, ft_dinv = P.build_vprime_table(7000.0, 0.0, 4e4, 1 << 16, spec.model)
    held = O.lookup(r, fine_t, ft_min, ft_dinv)
    netv = sur.vprime(r, 6950.0).astype(np.float64)
    err_held = np.max(np.abs(held - truth) / (r * 1e6))
    err_net = np.max(np.abs(netv - truth) / (r * 1e6))
    print(f"held-table error {err_held:.3e}, network error {err_net:.3e}")
    assert err_net < 0.2 * err_held


def test_grid_mode_matches_inline():
    """The per-step grid of the network (16385 points of the compressed coordinate, rebuilt at
    every step's T) gives the inline evaluation's fields: 10 cooling steps from a state that
    spans the barrier and the vacuum, within 1e-5 relative."""
    from tip import surrogate as S
    from tip.lattice import Lattice
    sur = S.Surrogate.load()
    rng = np.random.default_rng(31)
    shape = (16, 24, 64)
    rho = np.exp(rng.uniform(np.log(1.0), np.log(1.2e11), size=shape))
    th = rng.uniform(0, 2 * np.pi, size=shape)
    init = {"phi1": rho * np.cos(th), "phi2": rho * np.sin(th),
            "pi1": rng.normal(size=shape) * 1e2, "pi2": rng.normal(size=shape) * 1e2}
    kw = dict(dt=0.25, dx=1.0, eta_eff=1.0, mu=1000.0, noise_scale=300.0, seed=2)
    outs = []
    for grid in (0, 16385):
        lat = Lattice(shape, np.float32, complex_field=True)
        for k, v in init.items():
            lat.upload(k, v)
        lat.step_mlp(10, sur, T_start=7350.0, dT_per_step=-0.5, grid_points=grid, **kw)
        outs.append([lat.download(k).astype(np.float64) for k in ("phi1", "phi2", "pi1", "pi2")])
        lat.close()
    for a, b in zip(*outs):
        assert np.linalg.norm(a - b) <= 1e-5 * np.linalg.norm(b)
