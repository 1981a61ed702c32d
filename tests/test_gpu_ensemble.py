"""Batched ensemble on the GPU (-m gpu).

Bar: every execution mode (sequential per-instance launches, one batched launch per step,
hipGraph replay) and every launch geometry produce bit-identical fields, equal to the
single-lattice path with the same seed (which is itself checked against the oracle in
test_gpu_parity.py).  Graph replays are checked one by one with host work between them.
"""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

# z extents of every row mapping of the ensemble's row kernel: 64 and 8 (dense rows in
# 256-thread blocks), 96 and 24 (dense rows in 192-thread blocks), 70 and 65 (whole waves per
# row); the classes are interleaved so that the stepping order differs from the instance order
SHAPES = [(16, 16, 64), (24, 20, 70), (10, 12, 96), (8, 8, 8), (12, 33, 65), (6, 8, 24)]
SEEDS = (3, 11, 5, 0, 7, 9)


def _setup(seeds=SEEDS, geometry=None, shapes=SHAPES):
    from tip import potential as P
    from tip.ensemble import Ensemble
    ens = Ensemble(shapes, seeds=list(seeds))
    if geometry:
        ens.set_geometry(*geometry)
    rng = np.random.default_rng(1)
    init, tables, params = [], [], []
    p = P.ModelParams.real()
    for i, s in enumerate(shapes):
        phi = (rng.normal(size=s) * 2e3).astype(np.float32)
        pi = (rng.normal(size=s) * 3e2).astype(np.float32)
        ens.upload(i, "phi", phi)
        ens.upload(i, "pi", pi)
        T0 = 6770.0 + 150.0 * i
        t, tmin, dinv = P.build_vprime_table(T0, 0.0, 1e5, 2048, p)
        ens.set_table(i, t, tmin, dinv, odd=True)
        prm = dict(dt=0.25, dx=2.0, eta_eff=T0 / 1000.0, mu=1000.0,
                   noise_scale=P.noise_scale(T0 / 1000.0, T0, 0.25, 1000.0, 0.002))
        ens.set_params(i, **prm)
        init.append((phi, pi))
        tables.append((t, tmin, dinv))
        params.append(prm)
    return ens, init, tables, params


def _state(ens):
    return [(ens.download(i, "phi"), ens.download(i, "pi")) for i in range(ens.n)]


def _same(a, b):
    return all(np.array_equal(x[0], y[0]) and np.array_equal(x[1], y[1]) for x, y in zip(a, b))


@pytest.mark.parametrize("mode", ["sequential", "batched", "graph"])
def test_modes_match_single_lattice(mode):
    from tip.lattice import Lattice
    seeds = SEEDS
    ens, init, tables, params = _setup(seeds)
    ens.step(6, mode=mode)
    got = _state(ens)
    assert ens.step_count == 6
    for i, s in enumerate(SHAPES):
        lat = Lattice(s, np.float32)
        lat.upload("phi", init[i][0])
        lat.upload("pi", init[i][1])
        lat.set_table(*tables[i], odd=True)
        lat.step(6, seed=seeds[i], **params[i])
        assert np.array_equal(lat.download("phi"), got[i][0]), (mode, i)
        assert np.array_equal(lat.download("pi"), got[i][1]), (mode, i)
        lat.close()
    ens.close()


# (rows, planes, xcd, persistent blocks, kernel): flat tiles (2) and LDS row tiles (1); the
# default (auto) is row tiles with one y row per thread here (ny = 33 is odd)
@pytest.mark.parametrize("geometry", [(1, 3, 0, 0, 2), (2, 16, 1, 0, 2), (8, 4, 1, 16, 2),
                                      (4, 8, 1, 40, 2), (1, 3, 0, 0, 1), (1, 16, 1, 7, 1),
                                      (0, 5, 1, 0, 1), (1, 1, 0, 3, 1)])
def test_geometries_identical(geometry):
    ens_a, _, _, _ = _setup()
    ens_a.step(5, mode="batched")
    ref = _state(ens_a)
    ens_b, _, _, _ = _setup(geometry=geometry)
    ens_b.step(5, mode="graph")
    assert _same(_state(ens_b), ref)
    ens_a.close()
    ens_b.close()


# every ny divisible by 4, so two and four y rows per thread apply
SHAPES_TY = [(8, 16, 64), (6, 12, 96), (4, 8, 24), (6, 20, 70), (5, 8, 128), (3, 4, 192),
             (2, 8, 200)]


@pytest.mark.parametrize("ty", [2, 4])
@pytest.mark.parametrize("group_bytes", [0, 1])
def test_row_mappings_several_rows_per_thread(ty, group_bytes):
    """Row kernel with TY = 2 and 4 over dense rows (nz = 64, 128 in 256-thread blocks; 96, 24
    in 192-thread blocks, 192 filling one) and whole-wave rows (nz = 70, 200), one group or one
    group per instance, equal bit for bit to the single-lattice step."""
    from tip.lattice import Lattice
    seeds = (4, 8, 15, 16, 23, 42, 7)
    ens, init, tables, params = _setup(seeds, geometry=(ty, 3, 1, 0, 1), shapes=SHAPES_TY)
    ens.set_schedule(group_bytes)
    ens.step(2, mode="batched")
    ens.step(3, mode="graph")
    got = _state(ens)
    for i, s in enumerate(SHAPES_TY):
        lat = Lattice(s, np.float32)
        lat.upload("phi", init[i][0])
        lat.upload("pi", init[i][1])
        lat.set_table(*tables[i], odd=True)
        lat.step(5, seed=seeds[i], **params[i])
        assert np.array_equal(lat.download("phi"), got[i][0]), (ty, group_bytes, i)
        assert np.array_equal(lat.download("pi"), got[i][1]), (ty, group_bytes, i)
        lat.close()
    ens.close()


@pytest.mark.parametrize("group_bytes", [1, 40000, 1 << 30])
@pytest.mark.parametrize("concurrency", [1, 2])
def test_grouped_schedule_identical(group_bytes, concurrency):
    """Instances stepped group by group (each group through all steps of a call; with
    concurrency 2, two groups at once on two streams) give the same fields as one launch per
    step over all instances."""
    for mode in ("batched", "graph"):
        ens_b, _, _, _ = _setup()
        ens_b.set_schedule(group_bytes)
        ens_b.set_concurrency(concurrency)
        ens_b.step(2, mode=mode)
        ens_b.step(3, mode=mode)
        ens_b.step(70, mode=mode)  # a full 64-step graph chunk and a remainder
        ens_a2, _, _, _ = _setup()
        ens_a2.step(75, mode="batched")
        assert _same(_state(ens_b), _state(ens_a2)), (group_bytes, concurrency, mode)
        assert ens_b.step_count == 75
        ens_b.close()
        ens_a2.close()


def test_graph_replays_each_checked():
    ens_e, _, _, _ = _setup()
    ens_g, _, _, _ = _setup()
    for r in range(3):
        ens_e.step(4, mode="batched")
        ens_g.step(4, mode="graph")  # captured on r == 0, replayed afterwards
        a, b = _state(ens_e), _state(ens_g)
        assert _same(a, b), f"replay {r}"
        # host work between replays
        _ = [float(np.abs(x[0]).max()) for x in b]
    # odd step count flips the buffer parity: a second graph is captured for the other parity
    ens_e.step(3, mode="batched")
    ens_g.step(3, mode="graph")
    ens_e.step(3, mode="batched")
    ens_g.step(3, mode="graph")
    assert _same(_state(ens_e), _state(ens_g))
    assert ens_g.step_count == 18


def test_vacuum_counts_exact():
    ens, _, _, _ = _setup()
    ens.step(2, mode="graph")
    cnt = ens.vacuum_counts(1e3)
    st = _state(ens)
    for i in range(ens.n):
        assert int(cnt[i]) == int(np.sum(np.abs(st[i][0].astype(np.float64)) > 1e3))


def test_autotune_picks_and_keeps_results(tmp_path, monkeypatch):
    from tip import autotune
    monkeypatch.setattr(autotune, "CACHE_PATH", tmp_path / "cache.json")
    ens, _, _, _ = _setup()
    rec = autotune.tune_ensemble(ens, candidates=[(4, 8, 1, 0), (2, 4, 1, 0)], steps=2)
    assert rec["best"] in ([4, 8, 1, 0], [2, 4, 1, 0])
    assert (tmp_path / "cache.json").exists()
    rec2 = autotune.tune_ensemble(ens, candidates=[(8, 8, 1, 0)], steps=2)
    assert rec2["best"] == rec["best"]  # served from the cache


def test_joint_tuner_picks_a_pair_and_keeps_results(tmp_path, monkeypatch):
    from tip import autotune
    monkeypatch.setattr(autotune, "CACHE_PATH", tmp_path / "cache.json")
    geoms = [(1, 4, 1, 0, 1), (1, 8, 1, 0, 2), (1, 16, 1, 7, 1)]
    ens, _, _, _ = _setup()
    rec = autotune.tune_ensemble_joint(ens, geometries=geoms, group_sizes=(0, 1, 40000),
                                       top=2, steps=2)
    g, gb, conc = rec["best"]
    assert tuple(g) in geoms and gb in (0, 1, 40000) and conc in (1, 2)
    assert len(rec["table"]) == 2 * 3 * 2 and len(rec["geometries"]) == 3
    rec2 = autotune.tune_ensemble_joint(ens, geometries=[(1, 3, 0, 0, 1)],
                                        group_sizes=(0, 1, 40000), top=2, steps=2)
    assert rec2["best"] == rec["best"]  # served from the cache
    ens.close()
    # stepping under the tuned pair gives the untuned fields
    ens_a, _, _, _ = _setup()
    ens_a.step(4, mode="batched")
    ens_b, _, _, _ = _setup(geometry=tuple(g))
    ens_b.set_schedule(gb)
    ens_b.set_concurrency(conc)
    ens_b.step(4, mode="batched")
    assert _same(_state(ens_b), _state(ens_a))
    ens_a.close()
    ens_b.close()


def test_graph_after_table_param_and_geometry_changes():
    """Graph replays stay equal to batched launches bit for bit when tables (including a switch
    from odd to even tables, which changes the captured kernel variant), step parameters and
    the launch geometry change between calls, and for calls longer than one captured chunk."""
    from tip import potential as P
    ens_b, _, _, params = _setup()
    ens_g, _, _, _ = _setup()
    p = P.ModelParams.real()

    def both(fn):
        fn(ens_b)
        fn(ens_g)

    def step(k):
        ens_b.step(k, mode="batched")
        ens_g.step(k, mode="graph")
        assert _same(_state(ens_g), _state(ens_b)), k

    step(4)
    t, tmin, dinv = P.build_vprime_table(7300.0, 0.0, 1e5, 2048, p)
    both(lambda e: e.set_table(1, t, tmin, dinv, odd=True))
    prm = dict(params[2])
    prm["eta_eff"] = 9.5
    both(lambda e: e.set_params(2, **prm))
    step(4)
    # every instance to an even (signed) table: the odd/even kernel variant changes
    te, temin, tedinv = P.build_vprime_table(7000.0, -1e5, 1e5, 4097, p)
    for i in range(len(SHAPES)):
        both(lambda e: e.set_table(i, te, temin, tedinv, odd=False))
    step(4)
    both(lambda e: e.set_geometry(1, 5, 1, 0, 1))
    step(133)  # chunks of 64, 64 and 5 steps
    assert ens_g.step_count == ens_b.step_count == 145
    ens_b.close()
    ens_g.close()
