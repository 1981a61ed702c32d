"""GPU parity of the HIP kernels against the CPU oracle (run on an MI355X: -m gpu).

Bars (written into each test):
  * hash stream: x1, x2, u1, u2 bit-exact; z within 4e-15 relative (libm difference only)
  * float64 step: 100 RK2 steps at 64^3 with inline hash noise, field norms within 1e-12
    relative of the oracle (the reference's own float64 semantics)
  * float32 step: error against the float64 oracle stated and bounded
  * winding: per-plaquette int8 windings and string counts bit-exact
  * vacuum count, min/max: exact
"""
import math

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _tip():
    import tip
    from tip import kernels, kernels_complex
    from tip.device import Counter, DeviceArray
    return tip, kernels, kernels_complex, DeviceArray, Counter


def real_table(npts=8192):
    """8192 float64 points of V'(phi, T0 = 7000) on [0, 1.5 phi_true], odd."""
    from tip import potential as P
    p = P.ModelParams.real()
    vt = P.thermal_vev(7000.0, p)
    t, tmin, dinv = P.build_vprime_table(7000.0, 0.0, 1.5 * vt, npts, p)
    return t, tmin, dinv


def real_params():
    from tip import potential as P
    mu, dx_phys, dt_phys, T0 = 1000.0, 0.002, 2.5e-4, 7000.0
    dt, dx, eta = mu * dt_phys, mu * dx_phys, T0 / mu
    ns = P.noise_scale(eta, T0, dt, mu, dx_phys)
    return dict(dt=dt, dx=dx, eta_eff=eta, mu=mu, inv_a2=1.0, noise_scale=ns)


def rel(a, b):
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))


# ----------------------------------------------------------------------------------------
# hash RNG
# ----------------------------------------------------------------------------------------
def test_hash_bit_exact():
    tip, K, _, _, _ = _tip()
    edge = [0, 1, 2, 73856093, 0x9E3779B97F4A7C15, (1 << 63), (1 << 64) - 1, (1 << 32) - 1]
    seeds = np.concatenate([np.array(edge, np.uint64), O.site_seeds(1 << 20, 12345)])
    g = K.hash_rng_pairs(seeds)
    x1, x2, u1, u2 = O.hash_pairs(seeds)
    assert np.array_equal(g["x1"], x1) and np.array_equal(g["x2"], x2)
    assert np.array_equal(g["u1"], u1) and np.array_equal(g["u2"], u2)
    r = np.sqrt(-2.0 * np.log(u1))
    z = r * np.cos(6.283185307179586 * u2)
    zs = r * np.sin(6.283185307179586 * u2)
    assert np.max(np.abs(g["z"] - z) / np.maximum(np.abs(r), 1e-300)) < 4e-15
    assert np.max(np.abs(g["zs"] - zs) / np.maximum(np.abs(r), 1e-300)) < 4e-15
    assert K._hash_rng_pair(0) == (1e-12, float(u2[0]))


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_noise_field_matches_oracle(dtype):
    tip, K, _, DA, _ = _tip()
    shape = (16, 24, 40)
    for mult in (K.NOISE_MULT_PASS2, K.NOISE_MULT_PASS4):
        a1, a2 = O.noise_field(shape, 7, 1750.0, dtype=np.float64, pair=True, mult=mult)
        g1, g2 = np.empty(shape, dtype), np.empty(shape, dtype)
        K.generate_noise_field(g1, 1750.0, 7, mult=mult, noise2=g2)
        tol = 1e-14 if dtype == np.float64 else 2e-7
        assert np.max(np.abs(g1 - a1)) < tol * 1750.0 * 8
        assert np.max(np.abs(g2 - a2)) < tol * 1750.0 * 8


# ----------------------------------------------------------------------------------------
# real-field step
# ----------------------------------------------------------------------------------------
def _run_gpu_real(phi0, pi0, nsteps, prm, table, tmin, dinv, odd, dtype, seed=0):
    tip, K, _, DA, Counter = _tip()
    phi = DA.from_numpy(phi0.astype(dtype))
    pi = DA.from_numpy(pi0.astype(dtype))
    pt = DA(phi0.shape, dtype)
    qt = DA(phi0.shape, dtype)
    tab = DA.from_numpy(np.asarray(table, np.float64))
    for n in range(nsteps):
        K.rk2_step_table(phi, pi, prm["dt"], prm["dx"], prm["eta_eff"], prm["mu"], None, pt, qt,
                         prm["inv_a2"], tab, tmin, dinv, len(table), odd=odd,
                         noise_seed=O.noise_seed(n, seed), noise_scale=prm["noise_scale"])
    return phi.download(), pi.download()


def test_f64_100_steps_64cube_within_1e12():
    """64^3 real lattice, float64, 100 steps with hash noise: 1e-12 relative."""
    rng = np.random.default_rng(2024)
    shape = (64, 64, 64)
    phi0 = 0.01 * rng.normal(size=shape)
    pi0 = np.zeros(shape)
    prm = real_params()
    t, tmin, dinv = real_table()
    print(f"[f64 64^3] shape={shape} ns={prm['noise_scale']:.4g}", flush=True)
    a, b = O.run_real(phi0, pi0, 100, prm, t, tmin, dinv, odd=True)
    g, h = _run_gpu_real(phi0, pi0, 100, prm, t, tmin, dinv, True, np.float64)
    e_phi, e_pi = rel(g, a), rel(h, b)
    print(f"  rel err phi={e_phi:.3e} pi={e_pi:.3e}")
    assert e_phi <= 1e-12 and e_pi <= 1e-12
    # vacuum counts of the two runs agree exactly at a threshold inside the field's range
    esc = float(np.quantile(np.abs(a), 0.9))
    assert O.vacuum_count(g, esc) == O.vacuum_count(a, esc)


@pytest.mark.parametrize("shape", [(5, 7, 9), (3, 70, 130), (1, 1, 1), (2, 3, 64), (33, 5, 65),
                                   (2, 3, 1100)])
def test_f64_ragged_shapes_one_step_explicit_noise(shape):
    """Explicit noise array (the reference's calling form) on ragged shapes: 1e-14."""
    tip, K, _, DA, _ = _tip()
    rng = np.random.default_rng(sum(shape))
    phi = rng.normal(size=shape) * 3e3
    pi = rng.normal(size=shape) * 1e3
    noise = rng.normal(size=shape) * 500.0
    t = rng.normal(size=257) * 1e8
    tmin, dinv = -1e4, 256 / 2e4
    a, b = phi.copy(), pi.copy()
    O.rk2_step_table(a, b, 0.25, 2.0, 7.0, 1000.0, noise, np.empty_like(a), np.empty_like(b),
                     0.9, t, tmin, dinv, len(t))
    g, h = phi.copy(), pi.copy()
    K.rk2_step_table(g, h, 0.25, 2.0, 7.0, 1000.0, noise, np.empty_like(g), np.empty_like(h), 0.9,
                     t, tmin, dinv, len(t))
    assert rel(g, a) < 1e-14 and rel(h, b) < 1e-14


@pytest.mark.parametrize("shape", [(3, 70, 130), (4, 6, 1100), (5, 7, 9)])
def test_f64_hash_noise_ragged_shapes(shape):
    """Inline hash noise on ragged shapes: 3 steps within 1e-12 of the oracle."""
    rng = np.random.default_rng(sum(shape) + 1)
    phi0 = rng.normal(size=shape) * 3e3
    pi0 = rng.normal(size=shape) * 1e3
    t, tmin, dinv = real_table(1024)
    prm = real_params()
    a, b = O.run_real(phi0, pi0, 3, prm, t, tmin, dinv, odd=True, seed=9)
    g, h = _run_gpu_real(phi0, pi0, 3, prm, t, tmin, dinv, True, np.float64, seed=9)
    assert rel(g, a) <= 1e-12 and rel(h, b) <= 1e-12


@pytest.mark.parametrize("shape", [(128, 256, 256), (128, 258, 260)])
def test_f64_two_rows_per_thread_against_oracle(shape):
    """Lattices of 2^23 sites and more run the float64 row tiles with two y rows per thread
    (and, at nz = 260, five-wave rows ending in a wave of four active lanes): one step with
    inline hash noise from a random field within 1e-12 of the oracle."""
    rng = np.random.default_rng(shape[2])
    phi0 = rng.normal(size=shape) * 3e3
    pi0 = rng.normal(size=shape) * 1e3
    t, tmin, dinv = real_table(1024)
    prm = real_params()
    a, b = O.run_real(phi0, pi0, 1, prm, t, tmin, dinv, odd=True, seed=5)
    g, h = _run_gpu_real(phi0, pi0, 1, prm, t, tmin, dinv, True, np.float64, seed=5)
    assert rel(g, a) <= 1e-12 and rel(h, b) <= 1e-12


def test_f32_error_against_f64():
    """float32 kernel vs float64 oracle after 100 steps (64^3 real lattice): stated bound 1e-4."""
    rng = np.random.default_rng(7)
    shape = (64, 64, 64)
    phi0 = 0.01 * rng.normal(size=shape)
    pi0 = np.zeros(shape)
    prm = real_params()
    t, tmin, dinv = real_table()
    a, b = O.run_real(phi0, pi0, 100, prm, t, tmin, dinv, odd=True)
    g, h = _run_gpu_real(phi0, pi0, 100, prm, t, tmin, dinv, True, np.float32)
    e_phi, e_pi = rel(g.astype(np.float64), a), rel(h.astype(np.float64), b)
    print(f"[f32 vs f64, 100 steps] rel err phi={e_phi:.3e} pi={e_pi:.3e}")
    assert e_phi < 1e-4 and e_pi < 1e-4


def test_f32_long_run_against_f64_ensemble():
    """64^3 real lattice over 10^4 steps: the float32 path against the float64 path (itself within
    1e-12 of the oracle) for 4 seeds.  Past the first few hundred steps the fields of a chaotic
    Langevin run decorrelate whatever the precision, so the bar is on the true-vacuum fraction
    series: ensemble means within the ensembles' combined 1 sigma (plus one site)."""
    from tip.lattice import Lattice
    shape = (64, 64, 64)
    prm = real_params()
    t, tmin, dinv = real_table()
    kw = dict(dt=prm["dt"], dx=prm["dx"], eta_eff=prm["eta_eff"], mu=prm["mu"],
              noise_scale=prm["noise_scale"])
    marks = list(range(100, 2001, 100)) + list(range(3000, 10001, 1000))
    series = {np.float32: [], np.float64: []}
    err100, err_end = [], []
    for seed in range(4):
        phi0 = 0.01 * np.random.default_rng(100 + seed).normal(size=shape)
        lats = {}
        for dt_ in (np.float32, np.float64):
            lat = Lattice(shape, dt_)
            lat.upload("phi", phi0)
            lat.upload("pi", np.zeros(shape))
            lat.set_table(t, tmin, dinv, odd=True)
            lats[dt_] = lat
        fs = {np.float32: [], np.float64: []}
        done = 0
        for mk in marks:
            for dt_, lat in lats.items():
                lat.step(mk - done, seed=seed, step0=done, **kw)
                fs[dt_].append(lat.vacuum_fraction(1e4))
            if mk == 100:
                a = lats[np.float64].download("phi")
                err100.append(rel(lats[np.float32].download("phi").astype(np.float64), a))
            done = mk
        a = lats[np.float64].download("phi")
        err_end.append(rel(lats[np.float32].download("phi").astype(np.float64), a))
        for dt_, lat in lats.items():
            series[dt_].append(fs[dt_])
            lat.close()
    f32, f64 = np.array(series[np.float32]), np.array(series[np.float64])
    m32, m64 = f32.mean(0), f64.mean(0)
    s32, s64 = f32.std(0, ddof=1), f64.std(0, ddof=1)
    print(f"[f32 vs f64] field rel err after 100 steps {np.max(err100):.3e}, after 10^4 "
          f"{np.max(err_end):.3e}")
    for mk, a, b, c, d in zip(marks, m32, s32, m64, s64):
        print(f"  step {mk:6d}: f_true f32 {a:.5f} +- {b:.5f}   f64 {c:.5f} +- {d:.5f}")
    assert np.max(err100) < 1e-4
    tol = np.sqrt(s32 ** 2 + s64 ** 2) + 1.0 / np.prod(shape)
    assert np.all(np.abs(m32 - m64) <= tol)


def test_f32_matches_oracle_f32mixed_one_step():
    """float32 kernel vs the oracle's float32-mixed semantics, one step, explicit noise."""
    tip, K, _, DA, _ = _tip()
    rng = np.random.default_rng(3)
    shape = (16, 32, 64)
    phi = (rng.normal(size=shape) * 3e3).astype(np.float32)
    pi = (rng.normal(size=shape) * 1e3).astype(np.float32)
    noise = (rng.normal(size=shape) * 500.0).astype(np.float32)
    t = -1e6 * np.linspace(-1e5, 1e5, 1025)
    tmin, dinv = -1e5, 1024 / 2e5
    a, b = phi.copy(), pi.copy()
    O.rk2_step_table(a, b, 0.25, 2.0, 7.0, 1000.0, noise, np.empty_like(a), np.empty_like(b),
                     1.0, t, tmin, dinv, len(t))
    g, h = phi.copy(), pi.copy()
    K.rk2_step_table(g, h, 0.25, 2.0, 7.0, 1000.0, noise, np.empty_like(g), np.empty_like(h),
                     1.0, t, tmin, dinv, len(t))
    assert rel(g, a) < 1e-6 and rel(h, b) < 1e-6


def test_f32_dropin_two_streams_two_tables():
    """The float32 drop-in prepares its segment table in library scratch shared by every
    call on the device.  Calls on two streams with different tables, issued back to back,
    must each step with their own table: bit-identical to the same calls made one after the
    other on one stream."""
    from tip import _lib
    import ctypes as C
    tip, K, _, DA, _ = _tip()
    shape = (128, 128, 128)
    nx, ny, nz = shape
    prm = real_params()
    t1, tmin1, dinv1 = real_table()
    t2 = 1.5 * t1
    tabs = [(DA.from_numpy(t1), tmin1, dinv1), (DA.from_numpy(t2), tmin1, dinv1)]
    rng = np.random.default_rng(11)
    init = [((rng.normal(size=shape) * 3e3).astype(np.float32),
             (rng.normal(size=shape) * 1e3).astype(np.float32)) for _ in range(2)]

    def fields():
        return [(DA.from_numpy(a), DA.from_numpy(b), DA.empty(shape, np.float32),
                 DA.empty(shape, np.float32)) for a, b in init]

    def step(f, tab, seed, stream):
        phi, pi, pt, qt = f
        _lib.call("tip_rk2_step_table_f32", phi.ptr, pi.ptr, prm["dt"], prm["dx"],
                  prm["eta_eff"], prm["mu"], None, pt.ptr, qt.ptr, 1.0, tab[0].ptr, tab[1],
                  tab[2], len(t1), 1, nx, ny, nz, seed, prm["noise_scale"], None, 1e4, stream)

    streams = []
    for _ in range(2):
        h = C.c_void_p()
        _lib.call("tip_stream_create", C.byref(h))
        streams.append(h.value)
    try:
        ref = fields()
        for k in range(3):
            for i in range(2):
                step(ref[i], tabs[i], 100 + k, streams[0])
        _lib.call("tip_stream_synchronize", streams[0])
        got = fields()
        for k in range(3):
            for i in range(2):
                step(got[i], tabs[i], 100 + k, streams[i])
        for s_ in streams:
            _lib.call("tip_stream_synchronize", s_)
        for i in range(2):
            for a, b in zip(got[i][:2], ref[i][:2]):
                assert np.array_equal(a.numpy(), b.numpy())
    finally:
        for s_ in streams:
            _lib.call("tip_stream_destroy", s_)


def test_fused_vacuum_count_matches_standalone():
    tip, K, _, DA, Counter = _tip()
    rng = np.random.default_rng(11)
    shape = (32, 48, 64)
    phi = DA.from_numpy((rng.normal(size=shape) * 2e4).astype(np.float32))
    pi = DA.from_numpy((rng.normal(size=shape) * 1e3).astype(np.float32))
    pt, qt = DA(shape, np.float32), DA(shape, np.float32)
    t, tmin, dinv = -1e6 * np.linspace(-1e6, 1e6, 513), -1e6, 512 / 2e6
    c = Counter(1)
    K.rk2_step_table(phi, pi, 0.25, 2.0, 7.0, 1000.0, None, pt, qt, 1.0, t, tmin, dinv, len(t),
                     noise_seed=3, noise_scale=100.0, vac_count=c, vac_esc=1e4)
    fused = int(c.read()[0])
    host = phi.download()
    assert fused == int(np.sum(np.abs(host.astype(np.float64)) > 1e4))
    assert fused == K.vacuum_count(phi, 1e4)
    lo, hi = K.field_minmax(phi)
    assert lo == float(host.min()) and hi == float(host.max())


def test_fused_vacuum_count_float64_and_complex():
    """The counting instantiations of the float64 step and of the complex row-tile step: the
    fused count equals the standalone count of the new field (complex: up to the sites whose
    rho^2 lies within float32 rounding of esc^2, where the fused float32 test and the float64
    standalone test may differ)."""
    tip, K, KC, DA, Counter = _tip()
    rng = np.random.default_rng(12)
    shape = (16, 24, 64)
    t, tmin, dinv = -1e6 * np.linspace(-1e6, 1e6, 513), -1e6, 512 / 2e6
    phi = DA.from_numpy(rng.normal(size=shape) * 2e4)
    pi = DA.from_numpy(rng.normal(size=shape) * 1e3)
    pt, qt = DA(shape, np.float64), DA(shape, np.float64)
    c = Counter(1)
    K.rk2_step_table(phi, pi, 0.25, 2.0, 7.0, 1000.0, None, pt, qt, 1.0, t, tmin, dinv, len(t),
                     noise_seed=3, noise_scale=100.0, vac_count=c, vac_esc=1e4)
    assert int(c.read()[0]) == K.vacuum_count(phi, 1e4)

    from tip import potential as P
    ct, ctmin, cdinv = P.build_vprime_table(6770.0, 0.0, 1e5, 4096, P.ModelParams.complex())
    f = [DA.from_numpy((rng.normal(size=shape) * 1e4).astype(np.float32)) for _ in range(4)]
    tmp = [DA(shape, np.float32) for _ in range(4)]
    c = Counter(1)
    KC.rk2_step_table(*f, 0.25, 1.0, 1.0, 1000.0, None, None, *tmp, 1.0, ct, ctmin, cdinv,
                      len(ct), noise_seed=5, noise_scale=300.0, vac_count=c, vac_esc=1e4)
    fused = int(c.read()[0])
    a, b = f[0].download().astype(np.float64), f[1].download().astype(np.float64)
    r2 = a * a + b * b
    near = int(np.sum(np.abs(r2 - 1e8) < 1e8 * 2.5e-7))
    assert abs(fused - K.vacuum_count(f[0], 1e4, f[1])) <= near


# ----------------------------------------------------------------------------------------
# complex field
# ----------------------------------------------------------------------------------------
@pytest.mark.parametrize("zn", [(0.0, 0.0, 0.0), (4.0, 3e8, 1.0)])
def test_c64_20_steps(zn):
    tip, K, KC, DA, _ = _tip()
    rng = np.random.default_rng(5)
    shape = (24, 20, 36)
    f1, f2 = rng.normal(size=shape) * 1e3, rng.normal(size=shape) * 1e3
    p1, p2 = np.zeros(shape), np.zeros(shape)
    from tip import potential as P
    p = P.ModelParams.complex()
    t, tmin, dinv = P.build_vprime_table(6770.0, 0.0, 1e5, 4096, p)
    ns = P.noise_scale(1.0, 6770.0, 0.25, 1000.0, 1e-3)
    a = [f1.copy(), f2.copy(), p1.copy(), p2.copy()]
    tmp = [np.empty(shape) for _ in range(4)]
    for n in range(20):
        n1, n2 = O.noise_field(shape, O.noise_seed(n), ns, pair=True)
        O.rk2_step_table_complex(*a, 0.25, 1.0, 1.0, 1000.0, n1, n2, *tmp, 1.0, t, tmin, dinv,
                                 len(t), zn=zn)
    d = [DA.from_numpy(x) for x in (f1, f2, p1, p2)]
    dt_ = [DA(shape, np.float64) for _ in range(4)]
    for n in range(20):
        KC.rk2_step_table(*d, 0.25, 1.0, 1.0, 1000.0, None, None, *dt_, 1.0, t, tmin, dinv, len(t),
                          zn=zn, noise_seed=O.noise_seed(n), noise_scale=ns)
    g = [x.download() for x in d]
    for gi, ai in zip(g, a):
        assert rel(gi, ai) <= 1e-12


def test_c32_vs_oracle_mixed():
    tip, K, KC, DA, _ = _tip()
    rng = np.random.default_rng(6)
    shape = (16, 20, 64)
    f = [(rng.normal(size=shape) * 1e3).astype(np.float32) for _ in range(2)]
    p = [np.zeros(shape, np.float32) for _ in range(2)]
    n1, n2 = O.noise_field(shape, 1, 300.0, dtype=np.float32, pair=True)
    from tip import potential as P
    t, tmin, dinv = P.build_vprime_table(6770.0, 0.0, 1e5, 4096, P.ModelParams.complex())
    a = [f[0].copy(), f[1].copy(), p[0].copy(), p[1].copy()]
    O.rk2_step_table_complex(*a, 0.25, 1.0, 1.0, 1000.0, n1, n2,
                             *[np.empty(shape, np.float32) for _ in range(4)], 1.0, t, tmin, dinv,
                             len(t), zn=(3.0, 1e8, 1.0))
    g = [f[0].copy(), f[1].copy(), p[0].copy(), p[1].copy()]
    KC.rk2_step_table(*g, 0.25, 1.0, 1.0, 1000.0, n1, n2,
                      *[np.empty(shape, np.float32) for _ in range(4)], 1.0, t, tmin, dinv, len(t),
                      zn=(3.0, 1e8, 1.0))
    for gi, ai in zip(g, a):
        assert rel(gi, ai) < 1e-5


@pytest.mark.parametrize("zn_active", [1.0, 0.0])
def test_complex_lattice_geometries(zn_active):
    """Complex float32: every row-tile geometry gives identical fields (same per-site
    arithmetic); the flat tile (the generic formulation) agrees to float32 rounding.  Z_N on,
    and off (the row kernels with the Z_N term compiled out, including the per-block table
    choice of four-row tiles with shared halo rows)."""
    from tip import potential as P
    from tip.lattice import Lattice
    rng = np.random.default_rng(21)
    shape = (20, 36, 96)
    init = {k: (rng.normal(size=shape) * 2e3).astype(np.float32) for k in ("phi1", "phi2")}
    init.update({k: (rng.normal(size=shape) * 1e2).astype(np.float32) for k in ("pi1", "pi2")})
    t, tmin, dinv = P.build_vprime_table(6770.0, 0.0, 1e5, 4096, P.ModelParams.complex())
    kw = dict(dt=0.25, dx=1.0, eta_eff=1.0, mu=1000.0, noise_scale=300.0)
    outs = {}
    for geo in [(0, 0, 1, 0), (1, 3, 0, 1), (2, 7, 1, 1), (4, 40, 1, 1), (2, 1, 0, 1),
                (4, 8, 1, 2)]:
        lat = Lattice(shape, np.float32, complex_field=True)
        for k, v in init.items():
            lat.upload(k, v)
        lat.set_table(t, tmin, dinv)
        lat.set_zn(3.0, 1e8, zn_active)
        lat.set_geometry(*geo)
        lat.step(5, seed=4, **kw)
        outs[geo] = [lat.download(k) for k in ("phi1", "phi2", "pi1", "pi2")]
        lat.close()
    ref = outs[(0, 0, 1, 0)]
    for geo, o in outs.items():
        if geo[3] == 2:
            for a, b in zip(o, ref):
                assert rel(a.astype(np.float64), b.astype(np.float64)) < 1e-5, geo
        else:
            assert all(np.array_equal(a, b) for a, b in zip(o, ref)), geo


@pytest.mark.parametrize("zn_active", [1.0, 0.0])
@pytest.mark.parametrize("shape", [(6, 16, 250), (5, 12, 256), (4, 10, 128)])
def test_complex_shared_halo_rows(shape, zn_active):
    """Rows of two or four waves run two row groups per block that read their inner y-halo
    rows from each other's LDS staging (one, two and four rows per thread, x chunks that do
    and do not divide nx, ny that does and does not divide into pairs of tiles): identical
    fields to the unshared form's per-site arithmetic, the flat tile within float32
    rounding; Z_N on, and off (the Z_N term compiled out)."""
    from tip import potential as P
    from tip.lattice import Lattice, KERNEL_FLAT, KERNEL_ROWS
    rng = np.random.default_rng(sum(shape))
    init = {k: (rng.normal(size=shape) * 2e3).astype(np.float32) for k in ("phi1", "phi2")}
    init.update({k: (rng.normal(size=shape) * 1e2).astype(np.float32) for k in ("pi1", "pi2")})
    t, tmin, dinv = P.build_vprime_table(6770.0, 0.0, 1e5, 4096, P.ModelParams.complex())
    kw = dict(dt=0.25, dx=1.0, eta_eff=1.0, mu=1000.0, noise_scale=300.0)
    outs = {}
    geos = [(1, 0, 1, KERNEL_ROWS), (2, 3, 1, KERNEL_ROWS), (4, 0, 1, KERNEL_ROWS),
            (4, 5, 0, KERNEL_ROWS), (1, 0, 1, KERNEL_FLAT)]
    for geo in geos:
        if geo[3] == KERNEL_ROWS and shape[1] % geo[0]:
            continue
        lat = Lattice(shape, np.float32, complex_field=True)
        try:
            for k, v in init.items():
                lat.upload(k, v)
            lat.set_table(t, tmin, dinv)
            lat.set_zn(3.0, 1e8, zn_active)
            lat.set_geometry(*geo)
            lat.step(4, seed=8, step0=0, **kw)
            outs[geo] = [lat.download(k) for k in ("phi1", "phi2", "pi1", "pi2")]
        finally:
            lat.close()
    ref = outs[(1, 0, 1, KERNEL_ROWS)]
    for geo, o in outs.items():
        if geo[3] == KERNEL_FLAT:
            for a, b in zip(o, ref):
                assert rel(a.astype(np.float64), b.astype(np.float64)) < 1e-5, geo
        else:
            assert all(np.array_equal(a, b) for a, b in zip(o, ref)), geo


# ----------------------------------------------------------------------------------------
# winding, vacuum count
# ----------------------------------------------------------------------------------------
@pytest.mark.parametrize("shape", [(64, 64, 64), (7, 9, 70)])
def test_winding_bit_exact(shape):
    tip, K, KC, DA, _ = _tip()
    rng = np.random.default_rng(shape[2])
    p1 = rng.normal(size=shape).astype(np.float32)
    p2 = rng.normal(size=shape).astype(np.float32)
    w_ref, plaq_ref = O.winding(p1, p2)
    w, nstr, nplq, plaq = KC.compute_winding_number(p1, p2, plaquettes=True)
    assert np.array_equal(plaq, plaq_ref)
    assert np.array_equal(w, np.rint(w_ref).astype(np.float32))
    assert nstr == O.string_count(w_ref)
    assert nplq == int(np.count_nonzero(plaq_ref))
    # float64 inputs (the reference converts its float32 fields to float64)
    w2, nstr2, _, plaq2 = KC.compute_winding_number(p1.astype(np.float64), p2.astype(np.float64),
                                                    plaquettes=True)
    assert np.array_equal(plaq2, plaq_ref) and nstr2 == nstr


@pytest.mark.parametrize("shape", [(1, 1, 1), (1, 3, 5), (2, 2, 2), (3, 1, 65), (17, 5, 3),
                                   (33, 4, 130)])
def test_winding_small_and_odd_extents(shape):
    """x extents of 1, 2, 3 planes (the loads of plane x+2 wrap around the lattice while a
    block marches its chunk), chunks that end mid-lattice (17, 33 planes: 16-plane chunks plus
    a remainder), y extents of 1 (the y neighbour is the site itself) and z rows that are not
    a multiple of the 64-lane wave."""
    tip, K, KC, DA, _ = _tip()
    rng = np.random.default_rng(sum(shape) + 100)
    a1 = rng.normal(size=shape).astype(np.float32)
    a2 = rng.normal(size=shape).astype(np.float32)
    w_ref, plaq_ref = O.winding(a1, a2)
    for dtype in (np.float32, np.float64):  # float64 inputs hold the same (float32) values
        p1, p2 = a1.astype(dtype), a2.astype(dtype)
        w, nstr, nplq, plaq = KC.compute_winding_number(p1, p2, plaquettes=True)
        assert np.array_equal(plaq, plaq_ref), (shape, dtype)
        assert np.array_equal(w, np.rint(w_ref).astype(np.float32))
        assert nstr == O.string_count(w_ref) and nplq == int(np.count_nonzero(plaq_ref))


def test_winding_degenerate_edges():
    """Edges whose cross product is exactly zero: zero vectors with signed zeros (angle +-pi),
    the two sides of the negative real axis (theta = pi and -pi) and antiparallel pairs; the
    kernel falls back to the reference's atan2 expression there."""
    tip, K, KC, DA, _ = _tip()
    shape = (6, 8, 70)
    rng = np.random.default_rng(31)
    p1 = rng.choice(np.array([-2.0, -1.0, -0.0, 0.0, 1.0, 2.0], np.float32), size=shape)
    p2 = rng.choice(np.array([-1.0, -0.0, 0.0, 1.0], np.float32), size=shape)
    w_ref, plaq_ref = O.winding(p1, p2)
    w, nstr, nplq, plaq = KC.compute_winding_number(p1, p2, plaquettes=True)
    assert np.array_equal(plaq, plaq_ref)
    assert nstr == O.string_count(w_ref)


def test_winding_vortex_ring():
    """A vortex ring: winding lives on a closed loop of plaquettes; counts match."""
    tip, K, KC, DA, _ = _tip()
    n = 32
    i, j, k = np.meshgrid(*[np.arange(n, dtype=np.float64)] * 3, indexing="ij")
    c = (n - 1) / 2.0
    r = np.sqrt((i - c) ** 2 + (j - c) ** 2)
    # phase winds around the circle r = 8 in the (r, z) half plane
    p1 = (r - 8.0).astype(np.float32)
    p2 = (k - c).astype(np.float32)
    w_ref, plaq_ref = O.winding(p1, p2)
    w, nstr, nplq, plaq = KC.compute_winding_number(p1, p2, plaquettes=True)
    assert np.array_equal(plaq, plaq_ref)
    assert nstr == O.string_count(w_ref) and nstr > 0


def test_vacuum_counts_exact():
    tip, K, KC, DA, _ = _tip()
    rng = np.random.default_rng(9)
    for dtype in (np.float32, np.float64):
        phi = (rng.normal(size=(40, 50, 60)) * 1e4).astype(dtype)
        assert K.vacuum_count(phi, 1e4) == O.vacuum_count(phi, 1e4)
    a = (rng.normal(size=(40, 50, 60)) * 1e4).astype(np.float32)
    b = (rng.normal(size=(40, 50, 60)) * 1e4).astype(np.float32)
    assert K.vacuum_count(a, 1e4, b) == O.vacuum_count(a, 1e4, b)


@pytest.mark.parametrize("n", [1, 3, 4097, (1 << 22) + 3, 9_000_001])
def test_minmax_sizes_alignment_and_nan(n):
    """Field min/max (float32, float64) and the complex radius max: exact at several chunk
    counts, with aligned (16-byte loads) and misaligned (element loads) fields, and a NaN
    anywhere reaches both ends."""
    tip, K, KC, DA, Counter = _tip()
    from tip import _lib
    rng = np.random.default_rng(n + 7)
    out = DA((2,), np.float64)
    try:
        for dtype, name in ((np.float32, "tip_minmax_f32"), (np.float64, "tip_minmax_f64")):
            phi = (rng.normal(size=n + 1) * 1e4).astype(dtype)
            d = DA.from_numpy(phi)
            try:
                for off in (0, 1):
                    _lib.call(name, d.ptr + off * phi.itemsize, n, out.ptr, None)
                    r = out.download()
                    v = phi[off:off + n]
                    assert r[0] == float(v.min()) and r[1] == float(v.max()), (dtype, off)
                phi[n // 2] = np.nan
                d.upload(phi)
                _lib.call(name, d.ptr, n, out.ptr, None)
                r = out.download()
                assert np.isnan(r[0]) and np.isnan(r[1])
            finally:
                d.free()
        a = (rng.normal(size=n + 1) * 1e4).astype(np.float32)
        b = (rng.normal(size=n + 1) * 1e4).astype(np.float32)
        da, db = DA.from_numpy(a), DA.from_numpy(b)
        try:
            for oa, ob in ((0, 0), (1, 1), (0, 1)):
                rho = np.sqrt(a[oa:oa + n].astype(np.float64) ** 2
                              + b[ob:ob + n].astype(np.float64) ** 2)
                _lib.call("tip_rho_max_c32", da.ptr + 4 * oa, db.ptr + 4 * ob, n, out.ptr, None)
                assert out.download()[0] == float(rho.max())
        finally:
            da.free()
            db.free()
    finally:
        out.free()


@pytest.mark.parametrize("n", [0, 1, 3, 4097, 1 << 22, (1 << 22) + 3, 9_000_001])
def test_vacuum_counts_sizes_and_alignment(n):
    """Counts of 16-byte vector loads plus the element tail, several chunks per thread at the
    largest sizes, and fields that start off a 16-byte boundary (element loads)."""
    tip, K, KC, DA, Counter = _tip()
    from tip import _lib
    rng = np.random.default_rng(n)
    for dtype, name in ((np.float32, "tip_vacuum_count_f32"), (np.float64, "tip_vacuum_count_f64")):
        phi = (rng.normal(size=n + 1) * 1e4).astype(dtype)
        d = DA.from_numpy(phi)
        try:
            for off in (0, 1):
                c = Counter(1)
                _lib.call(name, d.ptr + off * phi.itemsize, n, 1e4, c.ptr, None)
                assert int(c.read()[0]) == O.vacuum_count(phi[off:off + n], 1e4), (dtype, off)
        finally:
            d.free()
    a = (rng.normal(size=n + 1) * 1e4).astype(np.float32)
    b = (rng.normal(size=n + 1) * 1e4).astype(np.float32)
    da, db = DA.from_numpy(a), DA.from_numpy(b)
    try:
        for oa, ob in ((0, 0), (1, 1), (0, 1)):
            c = Counter(1)
            _lib.call("tip_vacuum_count_c32", da.ptr + 4 * oa, db.ptr + 4 * ob, n, 1e4, c.ptr,
                      None)
            assert int(c.read()[0]) == O.vacuum_count(a[oa:oa + n], 1e4, b[ob:ob + n])
    finally:
        da.free()
        db.free()


# ----------------------------------------------------------------------------------------
# device-resident lattice: geometry independence and full-size properties
# ----------------------------------------------------------------------------------------
def test_lattice_geometries_bitwise_identical():
    """Every launch geometry computes the same per-site arithmetic: results are identical."""
    from tip.lattice import Lattice
    rng = np.random.default_rng(8)
    shape = (40, 36, 96)
    phi0 = (rng.normal(size=shape) * 1e3).astype(np.float32)
    pi0 = (rng.normal(size=shape) * 1e2).astype(np.float32)
    t, tmin, dinv = real_table(1024)
    prm = real_params()
    kw = dict(dt=prm["dt"], dx=prm["dx"], eta_eff=prm["eta_eff"], mu=prm["mu"],
              noise_scale=prm["noise_scale"])
    ref = None
    # (rows, planes, xcd, kernel): auto, flat tiles, LDS row tiles (nz = 96: half-idle wave)
    for by, cx, xcd, kind in [(0, 0, 1, 0), (1, 3, 0, 2), (2, 7, 1, 2), (4, 40, 0, 2),
                              (8, 1, 1, 2), (8, 16, 0, 2), (1, 5, 1, 1), (2, 3, 0, 1),
                              (4, 40, 1, 1), (2, 1, 1, 1)]:
        lat = Lattice(shape, np.float32)
        lat.upload("phi", phi0)
        lat.upload("pi", pi0)
        lat.set_table(t, tmin, dinv, odd=True)
        lat.set_geometry(by, cx, xcd, kind)
        lat.step(5, seed=3, **kw)
        out = (lat.download("phi"), lat.download("pi"))
        lat.close()
        if ref is None:
            ref = out
        else:
            assert np.array_equal(out[0], ref[0]) and np.array_equal(out[1], ref[1]), \
                (by, cx, xcd, kind)
    # and the drop-in entry point gives the same numbers
    from tip import kernels as K
    g, h = phi0.copy(), pi0.copy()
    for n in range(5):
        K.rk2_step_table(g, h, prm["dt"], prm["dx"], prm["eta_eff"], prm["mu"], None,
                         np.empty_like(g), np.empty_like(h), 1.0, t, tmin, dinv, len(t), odd=True,
                         noise_seed=O.noise_seed(n, 3), noise_scale=prm["noise_scale"])
    assert np.array_equal(g, ref[0]) and np.array_equal(h, ref[1])


@pytest.mark.parametrize("complex_field,shape", [(False, (3, 4, 1000)), (False, (3, 4, 450)),
                                                 (True, (3, 4, 512)), (True, (3, 4, 450)),
                                                 (True, (5, 6, 66))])
def test_row_tiles_wide_and_ragged_z(complex_field, shape):
    """Row tiles whose z rows span many waves (up to 16 real / 8 complex) and end in a wave
    with only a few active lanes (450 = 7 x 64 + 2, 66 = 64 + 2), whose idle lanes store past
    the plane's end: every row count gives identical fields, and the flat tile agrees to float32
    rounding."""
    from tip import potential as P
    from tip.lattice import Lattice, KERNEL_FLAT, KERNEL_ROWS
    rng = np.random.default_rng(sum(shape) + complex_field)
    names = ("phi1", "phi2", "pi1", "pi2") if complex_field else ("phi", "pi")
    init = {k: (rng.normal(size=shape) * (2e3 if k.startswith("phi") else 1e2)).astype(np.float32)
            for k in names}
    if complex_field:
        t, tmin, dinv = P.build_vprime_table(6770.0, 0.0, 1e5, 4096, P.ModelParams.complex())
        kw = dict(dt=0.25, dx=1.0, eta_eff=1.0, mu=1000.0, noise_scale=300.0)
    else:
        t, tmin, dinv = real_table(1024)
        prm = real_params()
        kw = dict(dt=prm["dt"], dx=prm["dx"], eta_eff=prm["eta_eff"], mu=prm["mu"],
                  noise_scale=prm["noise_scale"])
    outs = {}
    geos = [(1, 2, 1, KERNEL_ROWS), (2, 0, 0, KERNEL_ROWS), (1, 0, 1, KERNEL_FLAT)]
    for geo in geos:
        lat = Lattice(shape, np.float32, complex_field=complex_field)
        try:
            for k, v in init.items():
                lat.upload(k, v)
            lat.set_table(t, tmin, dinv, odd=not complex_field)
            lat.set_geometry(*geo)
            lat.step(4, seed=6, step0=0, **kw)
            outs[geo] = [lat.download(k) for k in names]
        finally:
            lat.close()
    a, b, f = (outs[g] for g in geos)
    assert all(np.array_equal(x, y) for x, y in zip(a, b))
    for x, y in zip(a, f):
        assert rel(x.astype(np.float64), y.astype(np.float64)) < 1e-5


@pytest.mark.parametrize("dtype,tol", [(np.float32, 1e-4), (np.float64, 1e-12)])
def test_full_size_256_uniform_mode_closed_form(dtype, tol):
    """256^3 (float32 and the float64 drop-in's kernel), no noise, linear table: a uniform
    field follows the exact RK2 matrix M = I + dt A + dt^2 A^2 / 2 over 100 steps."""
    from tip.lattice import Lattice
    n = 256
    c = -1.0e6
    x = np.linspace(-1e6, 1e6, 8193)
    t, tmin, dinv = c * x, -1e6, 8192 / 2e6
    lat = Lattice((n, n, n), dtype)
    lat.upload("phi", np.full((n, n, n), 300.0, dtype))
    lat.upload("pi", np.full((n, n, n), -50.0, dtype))
    lat.set_table(t, tmin, dinv)
    dt, eta, mu = 0.25, 7.0, 1000.0
    lat.step(100, dt=dt, dx=2.0, eta_eff=eta, mu=mu, noise_scale=0.0)
    A = np.array([[0.0, 1.0], [-c / mu ** 2, -eta]])
    M = np.eye(2) + dt * A + 0.5 * dt * dt * A @ A
    y = np.linalg.matrix_power(M, 100) @ np.array([300.0, -50.0])
    phi, pi = lat.download("phi"), lat.download("pi")
    assert np.all(phi == phi[0, 0, 0]) and np.all(pi == pi[0, 0, 0])  # stays uniform
    assert abs(phi[0, 0, 0] - y[0]) <= tol * abs(y[0])
    assert abs(pi[0, 0, 0] - y[1]) <= tol * abs(y[1])
    lo, hi = lat.minmax()
    assert lo == hi == float(phi[0, 0, 0])
    lat.close()


def test_full_size_256_noise_statistics():
    """256^3: one step from rest with no force isolates the noise: mean 0, var ns^2."""
    from tip.lattice import Lattice
    n = 256
    lat = Lattice((n, n, n), np.float32)
    lat.set_table(np.zeros(2), 0.0, 1.0)
    ns = 1750.0
    lat.step(1, dt=0.25, dx=2.0, eta_eff=7.0, mu=1000.0, noise_scale=ns, step0=42)
    pi = lat.download("pi").astype(np.float64).ravel()
    assert np.all(lat.download("phi") == 0.0)
    m, v = pi.mean(), pi.var()
    assert abs(m) < 5 * ns / math.sqrt(pi.size)
    assert abs(v / ns ** 2 - 1.0) < 5 * math.sqrt(2.0 / pi.size)
    # the same sites of the oracle stream (first 4096 sites), float32 transcendentals
    z = O.noise_field((1, 16, 256), 42, ns)
    np.testing.assert_allclose(pi[:4096], z.ravel(), rtol=0, atol=2e-6 * ns * 8)
    lat.close()


@pytest.mark.parametrize("integrator", ["rk2", "baoab", "overdamped"])
def test_full_size_256_default_geometry_bitwise(integrator):
    """256^3 real float32: the default launch (two rows per thread, ~4096 XCD-ordered blocks of
    8 planes) gives the same fields, bit for bit, as four rows per thread in 16-plane blocks
    and as one row per thread without the XCD order, on a random field with hash noise."""
    from tip.lattice import Lattice, KERNEL_AUTO, KERNEL_ROWS
    n = 256
    rng = np.random.default_rng(31)
    phi0 = (rng.normal(size=(n, n, n)) * 2e3).astype(np.float32)
    pi0 = (rng.normal(size=(n, n, n)) * 1e2).astype(np.float32)
    t, tmin, dinv = real_table(8192)
    prm = real_params()
    lat = Lattice((n, n, n), np.float32)
    lat.set_table(t, tmin, dinv, odd=True)
    outs = []
    try:
        for geo in [(0, 0, 1, KERNEL_AUTO), (4, 16, 1, KERNEL_ROWS), (1, 8, 0, KERNEL_ROWS)]:
            lat.set_geometry(*geo)
            lat.upload("phi", phi0)
            lat.upload("pi", pi0)
            if integrator == "rk2":
                lat.step(3, dt=prm["dt"], dx=prm["dx"], eta_eff=prm["eta_eff"], mu=prm["mu"],
                         noise_scale=prm["noise_scale"], seed=5, step0=0)
            elif integrator == "baoab":
                lat.step_baoab(3, dt=prm["dt"], dx=prm["dx"], gamma=prm["eta_eff"], mu=prm["mu"],
                               noise_scale=30.0, seed=5, step0=0)
            else:
                lat.step_overdamped(3, dt=prm["dt"], dx=prm["dx"], eta_eff=prm["eta_eff"],
                                    mu=prm["mu"], noise_scale=30.0, seed=5, step0=0)
            outs.append((lat.download("phi"), lat.download("pi")))
    finally:
        lat.close()
    for o in outs[1:]:
        assert np.array_equal(o[0], outs[0][0]) and np.array_equal(o[1], outs[0][1])
    assert np.isfinite(outs[0][0]).all() and float(np.abs(outs[0][0] - phi0).max()) > 0.0
