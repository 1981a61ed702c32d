"""Python face of the CPU oracle -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this module, and
only as the checker (or the timed CPU baseline).  It wraps oracle/libtip_oracle.so
(tip_oracle.c, a literal C restatement of the reference's rk2_step_table, _hash_rng_pair,
compute_winding_number and vacuum fraction) and holds a second, independent numpy
restatement of the hash RNG so the two can pin each other.

Reference anchors (paths relative to the reference checkout):
  simulation/latticeSimeRescale_numba.py:1875-1966   rk2_step_table (real)
  simulation/latticeSimComplex_numba.py:2067-2215    rk2_step_table (complex)
  simulation/latticeSimeRescale_numba.py:2322-2336   _hash_rng_pair
  simulation/latticeSimComplex_numba.py:2683-2744    compute_winding_number
  simulation/latticeSimeRescale_numba.py:1974-2192   baoab_step_table, overdamped_euler_step_table
  simulation/latticeSimComplex_numba.py:2223-2517    (complex forms of the two)
  postprocess/compute_lattice_Tc1.py:24              vacuum fraction
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
_LIB_PATH = HERE / "libtip_oracle.so"
_lib = None

_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_u64p = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")
_i8p = np.ctypeslib.ndpointer(dtype=np.int8, flags="C_CONTIGUOUS")
_d, _i64, _u64, _int = C.c_double, C.c_int64, C.c_uint64, C.c_int


class _NullableArray:
    """ctypes argtype accepting a contiguous numpy array of one dtype or None."""

    def __init__(self, dtype):
        self.dtype = np.dtype(dtype)

    def from_param(self, obj):
        if obj is None:
            return C.c_void_p(None)
        if not isinstance(obj, np.ndarray) or obj.dtype != self.dtype or not obj.flags.c_contiguous:
            raise TypeError(f"expected C-contiguous {self.dtype} array or None")
        return C.c_void_p(obj.ctypes.data)


_nf64 = _NullableArray(np.float64)
_nf32 = _NullableArray(np.float32)
_ni8 = _NullableArray(np.int8)


def lib():
    global _lib
    if _lib is None:
        if not _LIB_PATH.exists():
            raise RuntimeError(f"oracle library missing: {_LIB_PATH} (run oracle/build.py)")
        L = C.CDLL(str(_LIB_PATH))
        L.orc_hash_pairs.argtypes = [_u64p, _i64, _u64p, _u64p, _f64p, _f64p]
        L.orc_noise_field_f64.argtypes = [_f64p, _nf64, _i64, _i64, _i64, _u64, _i64, _d]
        L.orc_noise_field_f32.argtypes = [_f32p, _nf32, _i64, _i64, _i64, _u64, _i64, _d]
        L.orc_lookup.argtypes = [_f64p, _i64, _f64p, _f64p, _d, _d, _i64, _int]
        L.orc_rk2_step_table_f64.argtypes = [
            _f64p, _f64p, _d, _d, _d, _d, _nf64, _f64p, _f64p, _d, _f64p, _d, _d, _i64, _int,
            _i64, _i64, _i64]
        L.orc_rk2_step_table_f32mixed.argtypes = [
            _f32p, _f32p, _d, _d, _d, _d, _nf32, _f32p, _f32p, _d, _f64p, _d, _d, _i64, _int,
            _i64, _i64, _i64]
        L.orc_rk2_step_table_c64.argtypes = [
            _f64p, _f64p, _f64p, _f64p, _d, _d, _d, _d, _nf64, _nf64, _f64p, _f64p, _f64p, _f64p,
            _d, _f64p, _d, _d, _i64, _d, _d, _d, _i64, _i64, _i64]
        L.orc_rk2_step_table_c32mixed.argtypes = [
            _f32p, _f32p, _f32p, _f32p, _d, _d, _d, _d, _nf32, _nf32, _f32p, _f32p, _f32p, _f32p,
            _d, _f64p, _d, _d, _i64, _d, _d, _d, _i64, _i64, _i64]
        L.orc_langevin_f64.argtypes = [
            _int, _f64p, _nf64, _nf64, _nf64, _d, _d, _d, _d, _d, _d, _f64p, _d, _d, _i64, _int,
            _u64, _d, _i64, _i64, _i64]
        L.orc_langevin_c64.argtypes = [
            _int, _f64p, _f64p, _nf64, _nf64, _nf64, _nf64, _nf64, _nf64, _d, _d, _d, _d, _d, _d,
            _f64p, _d, _d, _i64, _d, _d, _d, _u64, _d, _i64, _i64, _i64]
        L.orc_complex_force.argtypes = [_f64p, _f64p, _i64, _f64p, _d, _d, _i64, _d, _d, _d,
                                        _f64p, _f64p]
        L.orc_cl_hash_ics.argtypes = [_f64p, _nf64, _i64, _d]
        L.orc_winding.argtypes = [_f64p, _f64p, _nf32, _ni8, _i64, _i64, _i64]
        L.orc_string_count.argtypes = [_f32p, _i64]
        L.orc_string_count.restype = _i64
        L.orc_vacuum_count_f64.argtypes = [_f64p, _i64, _d]
        L.orc_vacuum_count_f64.restype = _i64
        L.orc_vacuum_count_f32.argtypes = [_f32p, _i64, _d]
        L.orc_vacuum_count_f32.restype = _i64
        L.orc_vacuum_count_c32.argtypes = [_f32p, _f32p, _i64, _d]
        L.orc_vacuum_count_c32.restype = _i64
        L.orc_num_threads.restype = _int
        _lib = L
    return _lib


# ----------------------------------------------------------------------------------------
# Hash RNG
# ----------------------------------------------------------------------------------------
def hash_pairs(seeds):
    """C restatement: (x1, x2, u1, u2) for uint64 seeds."""
    seeds = np.ascontiguousarray(seeds, dtype=np.uint64)
    n = seeds.size
    x1 = np.empty(n, np.uint64)
    x2 = np.empty(n, np.uint64)
    u1 = np.empty(n, np.float64)
    u2 = np.empty(n, np.float64)
    lib().orc_hash_pairs(seeds.ravel(), n, x1, x2, u1, u2)
    return x1, x2, u1, u2


_M64 = (1 << 64) - 1


def hash_pair_py(seed: int):
    """Pure-Python restatement of _hash_rng_pair with explicit mod-2^64 arithmetic
    (simulation/latticeSimeRescale_numba.py:2322-2336), for a handful of seeds."""
    x = seed & _M64
    x = ((x ^ (x >> 30)) * 0xBF58476D1CE4E5B9) & _M64
    x = ((x ^ (x >> 27)) * 0x94D049BB133111EB) & _M64
    x = x ^ (x >> 31)
    x1 = x
    u1 = (x >> 11) * (1.0 / 9007199254740992.0)
    x = ((x * 0x2545F4914F6CDD1D) & _M64) ^ 0x9E3779B97F4A7C15
    x = ((x ^ (x >> 30)) * 0xBF58476D1CE4E5B9) & _M64
    x = x ^ (x >> 31)
    x2 = x
    u2 = (x >> 11) * (1.0 / 9007199254740992.0)
    if u1 < 1e-12:
        u1 = 1e-12
    return x1, x2, u1, u2


def hash_pairs_np(seeds):
    """numpy uint64 restatement (wrapping multiply), independent of the C text."""
    x = np.asarray(seeds, dtype=np.uint64).copy()
    with np.errstate(over="ignore"):
        x = (x ^ (x >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        x = (x ^ (x >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        x = x ^ (x >> np.uint64(31))
        x1 = x.copy()
        x = (x * np.uint64(0x2545F4914F6CDD1D)) ^ np.uint64(0x9E3779B97F4A7C15)
        x = (x ^ (x >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        x = x ^ (x >> np.uint64(31))
    x2 = x
    u1 = (x1 >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    u2 = (x2 >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    u1 = np.maximum(u1, 1e-12)
    return x1, x2, u1, u2


def site_seeds(n_sites: int, noise_seed: int, mult: int = 73856093):
    """uint64(site*mult) ^ noise_seed for site = 0..n_sites-1 (latticeSimeRescale_numba.py:2511-2516)."""
    s = np.arange(n_sites, dtype=np.int64)
    with np.errstate(over="ignore"):
        return (s * np.int64(mult)).astype(np.uint64) ^ np.uint64(noise_seed)


def noise_field(shape, noise_seed, scale, dtype=np.float64, pair=False, mult=73856093):
    """Hash noise field scale*z (and the sin partner if pair)."""
    nx, ny, nz = shape
    out1 = np.empty(shape, dtype)
    out2 = np.empty(shape, dtype) if pair else None
    fn = lib().orc_noise_field_f64 if dtype == np.float64 else lib().orc_noise_field_f32
    fn(out1, out2, nx, ny, nz, noise_seed, mult, scale)
    return (out1, out2) if pair else out1


def noise_seed(step: int, seed: int = 0) -> int:
    """step ^ (seed * golden) mod 2^64; seed 0 gives the reference's uint64(step)."""
    return (step ^ ((seed * 0x9E3779B97F4A7C15) & _M64)) & _M64


# ----------------------------------------------------------------------------------------
# Lookup and the step
# ----------------------------------------------------------------------------------------
def lookup(x, table, tmin, dinv, odd=False):
    x = np.ascontiguousarray(x, dtype=np.float64).ravel()
    out = np.empty_like(x)
    t = np.ascontiguousarray(table, dtype=np.float64)
    lib().orc_lookup(x, x.size, out, t, tmin, dinv, t.size, int(odd))
    return out


def rk2_step_table(phi, pi, dt, dx, eta_eff, mu, noise, phi_tmp, pi_tmp, inv_a2,
                   vp_table, vp_tmin, vp_dinv, vp_npts, odd=False):
    """Real-field step, in place, dtype of phi selects f64 or f32-mixed semantics."""
    nx, ny, nz = phi.shape
    t = np.ascontiguousarray(vp_table, dtype=np.float64)
    assert t.size == vp_npts
    if phi.dtype == np.float64:
        lib().orc_rk2_step_table_f64(phi, pi, dt, dx, eta_eff, mu, noise, phi_tmp, pi_tmp,
                                     inv_a2, t, vp_tmin, vp_dinv, vp_npts, int(odd), nx, ny, nz)
    elif phi.dtype == np.float32:
        lib().orc_rk2_step_table_f32mixed(phi, pi, dt, dx, eta_eff, mu, noise, phi_tmp, pi_tmp,
                                          inv_a2, t, vp_tmin, vp_dinv, vp_npts, int(odd),
                                          nx, ny, nz)
    else:
        raise TypeError(phi.dtype)


def complex_force(p1, p2, table, tmin, dinv, zn=(0.0, 0.0, 0.0)):
    """(dV1, dV2) of the complex step's force at points (p1, p2): V'(rho) phi_i / rho from
    the table plus the Z_N term (zn = (N, delta_V, active))."""
    p1 = np.ascontiguousarray(p1, dtype=np.float64).ravel()
    p2 = np.ascontiguousarray(p2, dtype=np.float64).ravel()
    t = np.ascontiguousarray(table, dtype=np.float64)
    a, b = np.empty_like(p1), np.empty_like(p1)
    lib().orc_complex_force(p1, p2, p1.size, t, tmin, dinv, t.size, float(zn[0]), float(zn[1]),
                            float(zn[2]), a, b)
    return a, b


def rk2_step_table_complex(phi1, phi2, pi1, pi2, dt, dx, eta_eff, mu, noise1, noise2,
                           phi1_tmp, phi2_tmp, pi1_tmp, pi2_tmp, inv_a2, vp_table, vp_tmin,
                           vp_dinv, vp_npts, zn=(0.0, 0.0, 0.0)):
    nx, ny, nz = phi1.shape
    t = np.ascontiguousarray(vp_table, dtype=np.float64)
    fn = (lib().orc_rk2_step_table_c64 if phi1.dtype == np.float64
          else lib().orc_rk2_step_table_c32mixed)
    fn(phi1, phi2, pi1, pi2, dt, dx, eta_eff, mu, noise1, noise2, phi1_tmp, phi2_tmp, pi1_tmp,
       pi2_tmp, inv_a2, t, vp_tmin, vp_dinv, vp_npts, float(zn[0]), float(zn[1]), float(zn[2]),
       nx, ny, nz)


# Overdamped Euler-Maruyama and BAOAB (tip_oracle.c: orc_langevin_f64 / orc_langevin_c64).
# op: 0 overdamped, 1 BAOAB open (B A O A), 2 fused (closing half-kick + B A O A), 3 close.
OVERDAMPED, BAOAB_OPEN, BAOAB_FUSED, BAOAB_CLOSE = 0, 1, 2, 3


def langevin(op, phi, pi, dt, dx, eta_or_gamma, mu, inv_a2, table, tmin, dinv, odd=False,
             noise_seed=0, noise_scale=0.0, inv_a2_prev=None):
    """One stage on float64 copies; returns (phi_out, pi_out) (None where the stage does not
    write)."""
    phi = np.ascontiguousarray(phi, dtype=np.float64)
    pi = None if pi is None else np.ascontiguousarray(pi, dtype=np.float64)
    t = np.ascontiguousarray(table, dtype=np.float64)
    nx, ny, nz = phi.shape
    po = np.empty_like(phi) if op != BAOAB_CLOSE else None
    qo = np.empty_like(phi) if op != OVERDAMPED else None
    lib().orc_langevin_f64(op, phi, pi, po, qo, dt, dx, eta_or_gamma, mu, inv_a2,
                           inv_a2 if inv_a2_prev is None else inv_a2_prev, t, tmin, dinv, t.size,
                           int(odd), noise_seed, noise_scale, nx, ny, nz)
    return po, qo


def langevin_complex(op, phi1, phi2, pi1, pi2, dt, dx, eta_or_gamma, mu, inv_a2, table, tmin,
                     dinv, zn=(0.0, 0.0, 0.0), noise_seed=0, noise_scale=0.0, inv_a2_prev=None):
    """Complex-field stage; returns (phi1_out, phi2_out, pi1_out, pi2_out)."""
    f = lambda a: None if a is None else np.ascontiguousarray(a, dtype=np.float64)
    phi1, phi2, pi1, pi2 = f(phi1), f(phi2), f(pi1), f(pi2)
    t = np.ascontiguousarray(table, dtype=np.float64)
    nx, ny, nz = phi1.shape
    mk = lambda c: np.empty_like(phi1) if c else None
    o = [mk(op != BAOAB_CLOSE), mk(op != BAOAB_CLOSE), mk(op != OVERDAMPED), mk(op != OVERDAMPED)]
    lib().orc_langevin_c64(op, phi1, phi2, pi1, pi2, *o, dt, dx, eta_or_gamma, mu, inv_a2,
                           inv_a2 if inv_a2_prev is None else inv_a2_prev, t, tmin, dinv, t.size,
                           float(zn[0]), float(zn[1]), float(zn[2]), noise_seed, noise_scale,
                           nx, ny, nz)
    return tuple(o)


def baoab_step(phi, pi, dt, dx, gamma, mu, inv_a2, table, tmin, dinv, odd=False, noise_seed=0,
               noise_scale=0.0):
    """One whole BAOAB step (open then close), float64 copies."""
    p1, q1 = langevin(BAOAB_OPEN, phi, pi, dt, dx, gamma, mu, inv_a2, table, tmin, dinv, odd,
                      noise_seed, noise_scale)
    _, q2 = langevin(BAOAB_CLOSE, p1, q1, dt, dx, gamma, mu, inv_a2, table, tmin, dinv, odd)
    return p1, q2


def run_real(phi, pi, nsteps, params, table, tmin, dinv, odd=False, seed=0, step0=0,
             noise_scale=None):
    """nsteps of the reference loop with hash noise: noise field for step n from
    noise_seed(n, seed), then rk2_step_table.  Returns (phi, pi) (copies)."""
    phi = phi.copy()
    pi = pi.copy()
    pt = np.empty_like(phi)
    qt = np.empty_like(pi)
    ns = params["noise_scale"] if noise_scale is None else noise_scale
    for n in range(step0, step0 + nsteps):
        noise = (noise_field(phi.shape, noise_seed(n, seed), ns, dtype=phi.dtype)
                 if ns != 0.0 else None)
        rk2_step_table(phi, pi, params["dt"], params["dx"], params["eta_eff"], params["mu"],
                       noise, pt, qt, params["inv_a2"], table, tmin, dinv, len(table), odd=odd)
    return phi, pi


# ----------------------------------------------------------------------------------------
# Observables
# ----------------------------------------------------------------------------------------
def winding(phi1, phi2):
    """Reference winding on float64 copies (latticeSimComplex_numba.py:4627-4629).
    Returns (w float32 site sums, plaq int8 [3, ...])."""
    p1 = np.ascontiguousarray(phi1, dtype=np.float64)
    p2 = np.ascontiguousarray(phi2, dtype=np.float64)
    nx, ny, nz = p1.shape
    w = np.empty(p1.shape, np.float32)
    plaq = np.empty((3,) + p1.shape, np.int8)
    lib().orc_winding(p1, p2, w, plaq, nx, ny, nz)
    return w, plaq


def string_count(w):
    w = np.ascontiguousarray(w, dtype=np.float32)
    return int(lib().orc_string_count(w.ravel(), w.size))


def vacuum_count(phi, esc, phi2=None):
    if phi2 is not None:
        a = np.ascontiguousarray(phi, dtype=np.float32).ravel()
        b = np.ascontiguousarray(phi2, dtype=np.float32).ravel()
        return int(lib().orc_vacuum_count_c32(a, b, a.size, esc))
    if phi.dtype == np.float64:
        return int(lib().orc_vacuum_count_f64(np.ascontiguousarray(phi).ravel(), phi.size, esc))
    a = np.ascontiguousarray(phi, dtype=np.float32).ravel()
    return int(lib().orc_vacuum_count_f32(a, a.size, esc))


def num_threads() -> int:
    return int(lib().orc_num_threads())


def cl_hash_ics(shape, amp, two=False):
    """The CosmoLattice port's hash initial conditions (tip_oracle.c orc_cl_hash_ics,
    cosmolattice_ext/models/thermal_inflation.h:174-222); float64 (f0, f1 or None)."""
    n = int(np.prod(shape))
    f0 = np.empty(n)
    f1 = np.empty(n) if two else None
    lib().orc_cl_hash_ics(f0, f1, n, float(amp))
    return f0.reshape(shape), (None if f1 is None else f1.reshape(shape))
