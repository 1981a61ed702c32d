"""The nucleation ensemble at its own workload on the GPU (-m gpu): 64 realisations
(32 x 64^3, 24 x 96^3, 8 x 128^3), seeds 1..64, T0 ~ U[6770, 7350] GeV, per-instance
8192-point float64 V' tables, float32 fields, the tuned launch geometry and step schedule.

Bars:
  * batched launches and hipGraph replay (a 16-step chunk replayed, then the remainder):
    all 64 instances bit-identical;
  * one instance of each size against the float64 oracle on the same hash stream
    (rk2_step_table, simulation/latticeSimeRescale_numba.py:1875-1966): relative field error
    below 1e-4, the float32 bar of the single-lattice parity tests;
  * per-instance vacuum counts exact (against numpy on the downloaded fields);
  * the nucleation bookkeeping of tip.driver.nucleation_ensemble (first report step with
    true-vacuum fraction > 0.5) reproduced from an independent batched run counted on the host.
"""
import numpy as np
import pytest

import bench
from oracle import oracle as O

pytestmark = pytest.mark.gpu

# the geometry and group size the bench's joint autotune picks on this workload: 4 y rows per
# thread, 16 x planes per work item, XCD-aware order, no persistent blocks, row kernel;
# about 320 MiB of field data per group
GEOMETRY = (4, 16, 1, 0, 1)
GROUP_BYTES = 320 << 20


def _ensemble(workload, tuned=True):
    from tip.ensemble import Ensemble
    shapes, seeds, _, inst = workload
    ens = Ensemble(shapes, seeds)
    for i, (t, tmin, dinv, prm) in enumerate(inst):
        ens.set_table(i, t, tmin, dinv, odd=True)
        ens.set_params(i, **prm)
    if tuned:
        ens.set_geometry(*GEOMETRY)
        ens.set_schedule(GROUP_BYTES)
    rng = np.random.default_rng(0)
    init = []
    for i, s in enumerate(shapes):
        phi = 0.01 * rng.normal(size=s)
        ens.upload(i, "phi", phi)
        ens.upload(i, "pi", np.zeros(s))
        init.append(phi.astype(np.float32).astype(np.float64))
    return ens, init


def rel(a, b):
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))


def test_workload_modes_oracle_and_counts():
    w = bench.ensemble_workload()
    shapes, seeds, T0s, inst = w
    assert len(shapes) == 64 and sum(int(np.prod(s)) for s in shapes) == 46399488
    assert all(len(x[0]) == 8192 for x in inst) and 6770 <= T0s.min() <= T0s.max() <= 7350
    nsteps = 20
    ens_b, init = _ensemble(w)
    ens_g, _ = _ensemble(w)
    ens_b.step(nsteps, mode="batched")
    ens_g.step(nsteps, mode="graph")
    assert ens_b.step_count == ens_g.step_count == nsteps
    for esc in (1e4, 30.0):
        counts = ens_b.vacuum_counts(esc)
        for i in range(64):
            phi_b = ens_b.download(i, "phi")
            if esc == 1e4:
                phi_g = ens_g.download(i, "phi")
                assert np.array_equal(phi_b, phi_g), f"phi of instance {i}"
                assert np.array_equal(ens_b.download(i, "pi"), ens_g.download(i, "pi")), i
            assert int(counts[i]) == int(np.sum(np.abs(phi_b.astype(np.float64)) > esc)), \
                (esc, i)
        print(f"esc {esc}: vacuum counts {int(counts.min())}..{int(counts.max())}", flush=True)
    for i in (0, 32, 56):  # one instance of each size: 64^3, 96^3, 128^3
        t, tmin, dinv, prm = inst[i]
        a, b = O.run_real(init[i], np.zeros(shapes[i]), nsteps, prm, t, tmin, dinv, odd=True,
                          seed=seeds[i])
        e_phi = rel(ens_b.download(i, "phi").astype(np.float64), a)
        e_pi = rel(ens_b.download(i, "pi").astype(np.float64), b)
        print(f"instance {i} {shapes[i]}: float32 vs float64 oracle after {nsteps} steps "
              f"phi {e_phi:.2e} pi {e_pi:.2e}", flush=True)
        assert e_phi < 1e-4 and e_pi < 1e-4
    ens_b.close()
    ens_g.close()


def test_nucleation_bookkeeping():
    from tip import driver
    steps, report = 300, 25
    res = driver.nucleation_ensemble(steps=steps, report=report, mode="graph")
    w = bench.ensemble_workload()
    ens, _ = _ensemble(w, tuned=False)
    sites = np.array([int(np.prod(s)) for s in w[0]])
    fr = []
    for _ in range(steps // report):
        ens.step(report, mode="batched")
        fr.append([np.sum(np.abs(ens.download(i, "phi").astype(np.float64)) > 1e4) / sites[i]
                   for i in range(64)])
    ens.close()
    fr = np.array(fr)
    assert np.array_equal(res["fractions"], fr)
    marks = report * np.arange(1, steps // report + 1)
    want = np.array([marks[np.argmax(fr[:, i] > 0.5)] if np.any(fr[:, i] > 0.5) else -1
                     for i in range(64)])
    assert np.array_equal(res["nucleation_step"], want)
    assert np.all(want > 0), "every realisation crosses f_true = 0.5 within 300 steps"
    assert np.allclose(res["T0"], w[2])
    print(f"nucleation steps {int(want.min())}..{int(want.max())} (median "
          f"{int(np.median(want))})", flush=True)
