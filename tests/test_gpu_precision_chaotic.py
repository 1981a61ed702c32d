"""float32 against float64 where the dynamics is chaotic (-m gpu).

Config 1's table (8192 points on [0, 1.5 phi_true], spacing ~1.8e11 GeV) does not resolve the
~1e3 GeV thermal barrier, so its runs roll to the true vacuum deterministically; here the
tables follow the drivers' own range rule (the field range +- max(range, 2e4 GeV), rebuilt
when T moves by 1e-4 or the field leaves the range, latticeSimeRescale_numba.py:3650-3672),
which resolves the barrier, and the runs sit where the outcome is decided by the noise.

  * real field, 64^3, T = 1220 GeV held fixed (just above the spinodal near 1150-1200 GeV):
    thermal nucleation out of the false vacuum at a random time; 8 seeds per precision,
    10^4 steps through the driver (float32 kernel vs the float64 drop-in kernel);
  * complex field, 64^3, linear cooling 7350 -> 6770 GeV over 1000 steps (the spinodal
    transition and the string network's decay); 4 seeds per precision through the
    rk2_step_table drop-in (c32 vs c64) with the same table rule for both.

Bars: the field's relative difference after 100 steps (before the trajectories decorrelate)
below 1e-4; the ensemble means of the true-vacuum fraction and (complex) the string count
within the two ensembles' combined 1 sigma (plus one site) at every report step.
"""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _within(a, b, extra):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    tol = np.sqrt(a.std(0, ddof=1) ** 2 + b.std(0, ddof=1) ** 2) + extra
    return np.abs(a.mean(0) - b.mean(0)) <= tol


def test_real_nucleation_f32_vs_f64():
    from tip import driver as D
    series = {"f32": [], "f64": []}
    err100 = []
    for seed in range(8):
        fields = {}
        for prec in ("f32", "f64"):
            cfg = D.RunConfig(field="real", Nx=64, Ny=64, Nz=64, Nt=100, T0=1220.0,
                              precision=prec, report=100, seed=seed, quiet=True, divergence=0.0)
            run = D.Run(cfg)
            run.run()
            fields[prec] = run.lat.download("phi").astype(np.float64)
            run.close()
            # the series: one uninterrupted run per precision (its first 100 steps are these)
            cfg = D.RunConfig(field="real", Nx=64, Ny=64, Nz=64, Nt=10000, T0=1220.0,
                              precision=prec, report=250, seed=seed, quiet=True, divergence=0.0)
            run = D.Run(cfg)
            ser = run.run()
            run.close()
            series[prec].append(ser["f_true"])
        a = fields["f64"]
        err100.append(np.linalg.norm(fields["f32"] - a) / np.linalg.norm(a))
    f32, f64 = np.array(series["f32"]), np.array(series["f64"])
    marks = 250 * np.arange(1, f32.shape[1] + 1)
    nuc = {k: [int(marks[np.argmax(np.asarray(v) > 0.5)]) if np.max(v) > 0.5 else -1
               for v in series[k]] for k in series}
    print(f"real 64^3 at 1220 GeV: field rel diff after 100 steps {max(err100):.2e}; "
          f"nucleation steps f32 {nuc['f32']} f64 {nuc['f64']}", flush=True)
    for k, (a, sa, b, sb) in enumerate(zip(f32.mean(0), f32.std(0, ddof=1), f64.mean(0),
                                           f64.std(0, ddof=1))):
        if k % 8 == 7:
            print(f"  step {marks[k]:6d}: f_true f32 {a:.3f} +- {sa:.3f}   f64 {b:.3f} +- {sb:.3f}")
    assert max(err100) < 1e-4
    assert np.all(_within(f32, f64, 1.0 / 64 ** 3))
    # the regime is stochastic: the ensemble is neither all-nucleated nor all-stuck at the end
    assert 0 < sum(n > 0 for n in nuc["f64"]) and sum(n > 0 for n in nuc["f32"]) > 0


def _complex_run(dtype, seed, steps=1000, report=100, n=64):
    from tip import driver as D
    from tip import kernels_complex as KC
    from tip import potential as P
    from tip import vprime_device as VD
    from tip.device import DeviceArray
    cfg = D.RunConfig(field="complex", Nx=n, Ny=n, Nz=n, Nt=steps, T0=7350.0, cool_to=6770.0,
                      seed=seed, integrator="rk2")
    f0 = D.initial_fields(cfg)
    d = [DeviceArray.from_numpy(f0[k].astype(dtype)) for k in ("phi1", "phi2", "pi1", "pi2")]
    tmp = [DeviceArray(d[0].shape, dtype) for _ in range(4)]
    mu, dt, dx, eta = D.MU, D.MU * cfg.dt_phys, D.MU * cfg.dx_phys, cfg.eta_phys / D.MU
    model = D.model_for(cfg)
    table = DeviceArray((cfg.table_size,), np.float64)
    tT = hi = None
    tmin = dinv = 0.0
    f_ser, s_ser, out100 = [], [], None
    for k in range(steps):
        T = cfg.T0 + (cfg.cool_to - cfg.T0) * k / steps
        need = tT is None or abs(T - tT) / T > 1e-4
        if k % 5 == 0 or need:
            a, b = d[0].download().astype(np.float64), d[1].download().astype(np.float64)
            rmax = float(np.sqrt(a * a + b * b).max())
            if need or rmax > hi:
                margin = max(rmax, 2e4)
                hi = rmax + margin
                _, tmin, dinv = VD.build_vprime_table(T, 0.0, hi, cfg.table_size, model,
                                                      out=table)
                tT = T
        ns = P.noise_scale(eta, T, dt, mu, cfg.dx_phys)
        KC.rk2_step_table(*d, dt, dx, eta, mu, None, None, *tmp, 1.0, table, tmin, dinv,
                          cfg.table_size, zn=(0.0, 0.0, 0.0), noise_seed=O.noise_seed(k, seed),
                          noise_scale=ns)
        if k + 1 == 100:
            out100 = [x.download().astype(np.float64) for x in d[:2]]
        if (k + 1) % report == 0:
            a, b = d[0].download().astype(np.float64), d[1].download().astype(np.float64)
            f_ser.append(np.count_nonzero(a * a + b * b > 1e8) / n ** 3)
            s_ser.append(KC.compute_winding_number(d[0], d[1])[1])
    for x in d + tmp + [table]:
        x.free()
    return np.array(f_ser), np.array(s_ser), out100


def test_complex_cooling_f32_vs_f64():
    res = {np.float32: [], np.float64: []}
    err100 = []
    for seed in range(4):
        outs = {}
        for dt_ in (np.float32, np.float64):
            f, s, o = _complex_run(dt_, seed)
            res[dt_].append((f, s))
            outs[dt_] = o
        err100.append(max(np.linalg.norm(a - b) / np.linalg.norm(b)
                          for a, b in zip(outs[np.float32], outs[np.float64])))
    F32 = np.array([r[0] for r in res[np.float32]])
    F64 = np.array([r[0] for r in res[np.float64]])
    S32 = np.array([r[1] for r in res[np.float32]])
    S64 = np.array([r[1] for r in res[np.float64]])
    print(f"complex 64^3 cooling: field rel diff after 100 steps {max(err100):.2e}", flush=True)
    print("  f_true f32", np.round(F32.mean(0), 3).tolist(), "f64", np.round(F64.mean(0), 3).tolist())
    print("  strings f32", np.round(S32.mean(0)).astype(int).tolist(), "f64",
          np.round(S64.mean(0)).astype(int).tolist(), flush=True)
    assert max(err100) < 1e-4
    assert np.all(_within(F32, F64, 1.0 / 64 ** 3))
    assert np.all(_within(S32, S64, 1.0))
    assert F64.mean(0)[0] == 0.0 and F64.mean(0)[-1] == 1.0
