"""Drop-in replacements for the real-field kernels of simulation/latticeSimeRescale_numba.py.

Same names, same positional arguments, same in-place semantics as the reference's Numba
functions; the work runs in libtip's gfx950 kernels.

  rk2_step_table(phi, pi, dt, dx, eta_eff, mu, noise, phi_tmp, pi_tmp, inv_a2,
                 vp_table, vp_tmin, vp_dinv, vp_npts)           :1875-1966
  overdamped_euler_step_table(phi, dt, dx, eta_eff, mu, inv_a2,
                 vp_table, vp_tmin, vp_dinv, vp_npts, noise_scale, seed)   :2115-2192
  baoab_step_table(phi, pi, dt, dx, gamma, mu, inv_a2,
                 vp_table, vp_tmin, vp_dinv, vp_npts, noise_scale, seed)   :1974-2112
  generate_noise_field(noise, scale, seed)                        :1348-1361
  _hash_rng_pair(seed) -> (u1, u2)                                :2322-2336

Arrays may be numpy arrays (host: each call uploads, steps and downloads -- the parity
form) or tip.device.DeviceArray (device-resident, no transfers -- the performance form).
Dtype and shape errors raise TypeError/ValueError where the reference's dispatcher raises
TypingError.

Noise: the reference adds a pre-generated noise field.  Passing noise=None selects the
library's inline hash noise: keyword noise_scale (the reference's _ns_val,
:3714-3718) and noise_seed (uint64(step) in the reference's fused kernels, :3738).
"""
from __future__ import annotations

import ctypes as C
from typing import Optional

import numpy as np

from . import _lib
from .device import DeviceArray

NOISE_MULT_PASS2 = 73856093  # latticeSimeRescale_numba.py:2514
NOISE_MULT_PASS4 = 19349669  # latticeSimeRescale_numba.py:2639


class _Staging:
    """Device buffers reused across host-array calls of the same shape and dtype."""

    def __init__(self):
        self._bufs = {}

    def get(self, key, shape, dtype) -> DeviceArray:
        k = (key, tuple(shape), np.dtype(dtype).str)
        a = self._bufs.get(k)
        if a is None:
            a = DeviceArray(shape, dtype)
            self._bufs[k] = a
        return a

    def clear(self):
        for a in self._bufs.values():
            a.free()
        self._bufs.clear()


_staging = _Staging()


def _is_dev(x) -> bool:
    return isinstance(x, DeviceArray)


def _check_field(name, x, shape=None, dtype=None):
    if x is None:
        raise TypeError(f"{name} must be an array")
    if not (_is_dev(x) or isinstance(x, np.ndarray)):
        raise TypeError(f"{name}: expected numpy array or DeviceArray, got {type(x).__name__}")
    if len(x.shape) != 3:
        raise TypeError(f"{name}: expected a 3-D array, got shape {x.shape}")
    if dtype is not None and np.dtype(x.dtype) != np.dtype(dtype):
        raise TypeError(f"{name}: dtype {x.dtype} != {np.dtype(dtype)}")
    if shape is not None and tuple(x.shape) != tuple(shape):
        raise ValueError(f"{name}: shape {x.shape} != {shape}")
    if isinstance(x, np.ndarray) and not x.flags.c_contiguous:
        raise ValueError(f"{name}: array must be C-contiguous")


def _table_dev(vp_table, vp_npts):
    if _is_dev(vp_table):
        if vp_table.dtype != np.float64 or vp_table.shape[0] != vp_npts:
            raise ValueError("vp_table must be float64 with vp_npts entries")
        return vp_table
    t = np.ascontiguousarray(vp_table, dtype=np.float64)
    if t.ndim != 1 or t.size != int(vp_npts):
        raise ValueError(f"vp_table has {t.size} entries, vp_npts={vp_npts}")
    if vp_npts < 2:
        raise ValueError("vp_npts must be >= 2")
    d = _staging.get("table", t.shape, np.float64)
    d.upload(t)
    return d


def rk2_step_table(phi, pi, dt, dx, eta_eff, mu, noise, phi_tmp, pi_tmp, inv_a2, vp_table,
                   vp_tmin, vp_dinv, vp_npts, *, noise_seed: int = 0, noise_scale: float = 0.0,
                   odd: bool = False, vac_count=None, vac_esc: float = 1e4, stream=None):
    """One RK2 Langevin step of the real field, in place on (phi, pi).

    Reference: rk2_step_table, simulation/latticeSimeRescale_numba.py:1875-1966.
    float64 arrays run the float64 drop-in (reference operation order); float32 arrays the
    float32 kernel.
    """
    _check_field("phi", phi)
    shape, dtype = tuple(phi.shape), np.dtype(phi.dtype)
    if dtype not in (np.float32, np.float64):
        raise TypeError(f"phi: unsupported dtype {dtype}")
    for nm, a in (("pi", pi), ("phi_tmp", phi_tmp), ("pi_tmp", pi_tmp)):
        _check_field(nm, a, shape, dtype)
    if noise is not None:
        _check_field("noise", noise, shape, dtype)
    table = _table_dev(vp_table, int(vp_npts))
    nx, ny, nz = shape
    host = not _is_dev(phi)
    if host and any(_is_dev(a) for a in (pi, phi_tmp, pi_tmp, noise) if a is not None):
        raise TypeError("mix of host and device arrays")
    if host:
        d_phi = _staging.get("phi", shape, dtype)
        d_pi = _staging.get("pi", shape, dtype)
        d_pt = _staging.get("phi_tmp", shape, dtype)
        d_qt = _staging.get("pi_tmp", shape, dtype)
        d_phi.upload(phi)
        d_pi.upload(pi)
        d_noise = None
        if noise is not None:
            d_noise = _staging.get("noise", shape, dtype)
            d_noise.upload(noise)
    else:
        d_phi, d_pi, d_pt, d_qt, d_noise = phi, pi, phi_tmp, pi_tmp, noise
    name = "tip_rk2_step_table_f64" if dtype == np.float64 else "tip_rk2_step_table_f32"
    _lib.call(name, d_phi.ptr, d_pi.ptr, float(dt), float(dx), float(eta_eff), float(mu),
              d_noise.ptr if d_noise is not None else None, d_pt.ptr, d_qt.ptr, float(inv_a2),
              table.ptr, float(vp_tmin), float(vp_dinv), int(vp_npts), int(bool(odd)), nx, ny,
              nz, int(noise_seed) & 0xFFFFFFFFFFFFFFFF, float(noise_scale),
              vac_count.ptr if vac_count is not None else None, float(vac_esc), stream)
    if host:
        d_phi.download(phi)
        d_pi.download(pi)


BAOAB_OPEN, BAOAB_FUSED, BAOAB_CLOSE = 1, 2, 3  # TIP_BAOAB_* of include/tip.h


def _field_set(names, arrays, dtype):
    """Device buffers for a call: the arrays themselves when on the device, otherwise staging
    buffers holding an upload of each (host arrays).  Returns (devices, host?)."""
    host = not _is_dev(arrays[0])
    if host and any(_is_dev(a) for a in arrays):
        raise TypeError("mix of host and device arrays")
    if not host:
        return list(arrays), False
    devs = []
    for nm, a in zip(names, arrays):
        d = _staging.get(nm, a.shape, dtype)
        d.upload(a)
        devs.append(d)
    return devs, True


def _write_back(host, dst, src):
    """dst <- src after a call: a download into host arrays, a device copy otherwise."""
    for a, d in zip(dst, src):
        if host:
            d.download(a)
        else:
            a.copy_from(d)


def overdamped_euler_step_table(phi, dt, dx, eta_eff, mu, inv_a2, vp_table, vp_tmin, vp_dinv,
                                vp_npts, noise_scale, seed, *, odd: bool = False, stream=None):
    """One overdamped Euler-Maruyama step of the real field, in place on phi.

    Reference: overdamped_euler_step_table, simulation/latticeSimeRescale_numba.py:2115-2192,
    evaluated on one consistent field (the reference's in-place loop lets part of each
    Laplacian see advanced neighbours).  seed is the reference's per-call seed (the step
    index); the noise is the hash stream of noise_seed = seed.
    """
    _check_field("phi", phi)
    shape, dtype = tuple(phi.shape), np.dtype(phi.dtype)
    if dtype not in (np.float32, np.float64):
        raise TypeError(f"phi: unsupported dtype {dtype}")
    table = _table_dev(vp_table, int(vp_npts))
    (d_phi,), host = _field_set(["od_phi"], [phi], dtype)
    out = _staging.get("od_out", shape, dtype)
    name = "tip_overdamped_f64" if dtype == np.float64 else "tip_overdamped_f32"
    _lib.call(name, d_phi.ptr, out.ptr, float(dt), float(dx), float(eta_eff), float(mu),
              float(inv_a2), table.ptr, float(vp_tmin), float(vp_dinv), int(vp_npts),
              int(bool(odd)), *shape, int(seed) & 0xFFFFFFFFFFFFFFFF, float(noise_scale), stream)
    _write_back(host, [phi], [out])


def baoab_step_table(phi, pi, dt, dx, gamma, mu, inv_a2, vp_table, vp_tmin, vp_dinv, vp_npts,
                     noise_scale, seed, *, odd: bool = False, stream=None):
    """One BAOAB Langevin step of the real field, in place on (phi, pi).

    Reference: baoab_step_table, simulation/latticeSimeRescale_numba.py:1974-2112 (two lattice
    passes: B A O A, then the closing B on the advanced field), with every force evaluated on
    one consistent field.  noise_scale is the reference's BAOAB amplitude sqrt(T / (mu^2
    dx^3)); c2 = sqrt(1 - exp(-2 gamma dt)) noise_scale.
    """
    _check_field("phi", phi)
    shape, dtype = tuple(phi.shape), np.dtype(phi.dtype)
    if dtype not in (np.float32, np.float64):
        raise TypeError(f"phi: unsupported dtype {dtype}")
    _check_field("pi", pi, shape, dtype)
    table = _table_dev(vp_table, int(vp_npts))
    (d_phi, d_pi), host = _field_set(["ba_phi", "ba_pi"], [phi, pi], dtype)
    po = _staging.get("ba_phi_out", shape, dtype)
    qo = _staging.get("ba_pi_out", shape, dtype)
    name = "tip_baoab_f64" if dtype == np.float64 else "tip_baoab_f32"
    common = (float(dt), float(dx), float(gamma), float(mu), float(inv_a2), float(inv_a2),
              table.ptr, float(vp_tmin), float(vp_dinv), int(vp_npts), int(bool(odd)), *shape)
    _lib.call(name, BAOAB_OPEN, d_phi.ptr, d_pi.ptr, po.ptr, qo.ptr, *common,
              int(seed) & 0xFFFFFFFFFFFFFFFF, float(noise_scale), stream)
    _lib.call(name, BAOAB_CLOSE, po.ptr, qo.ptr, None, qo.ptr, *common, 0, 0.0, stream)
    _write_back(host, [phi, pi], [po, qo])


def generate_noise_field(noise, scale, seed, *, mult: int = NOISE_MULT_PASS2, noise2=None,
                         stream=None):
    """Fill `noise` with scale * z of the hash stream for noise_seed = seed.

    Replaces generate_noise_field (latticeSimeRescale_numba.py:1348-1361), whose
    per-thread np.random stream is not reproducible, with the deterministic stream of the
    reference's fused kernels (:2511-2522).  noise2 (optional) receives the sin partner.
    """
    _check_field("noise", noise)
    shape, dtype = tuple(noise.shape), np.dtype(noise.dtype)
    if noise2 is not None:
        _check_field("noise2", noise2, shape, dtype)
    nx, ny, nz = shape
    name = "tip_noise_field_f64" if dtype == np.float64 else "tip_noise_field_f32"
    if _is_dev(noise):
        _lib.call(name, noise.ptr, noise2.ptr if noise2 is not None else None, nx, ny, nz,
                  int(seed) & 0xFFFFFFFFFFFFFFFF, int(mult), float(scale), stream)
        return
    d1 = _staging.get("noise", shape, dtype)
    d2 = _staging.get("noise2", shape, dtype) if noise2 is not None else None
    _lib.call(name, d1.ptr, d2.ptr if d2 is not None else None, nx, ny, nz,
              int(seed) & 0xFFFFFFFFFFFFFFFF, int(mult), float(scale), stream)
    d1.download(noise)
    if d2 is not None:
        d2.download(noise2)


def hash_rng_pairs(seeds: np.ndarray):
    """Vectorised _hash_rng_pair on the device.

    Returns dict with x1, x2 (uint64 mixer states), u1, u2 (float64), z, zs (float64
    Box-Muller cos/sin components).
    """
    s = np.ascontiguousarray(seeds, dtype=np.uint64).ravel()
    n = s.size
    out = {}
    if n == 0:
        for k, dt in (("x1", np.uint64), ("x2", np.uint64), ("u1", np.float64),
                      ("u2", np.float64), ("z", np.float64), ("zs", np.float64)):
            out[k] = np.empty(0, dt)
        return out
    d_s = DeviceArray.from_numpy(s)
    bufs = {k: DeviceArray((n,), dt) for k, dt in (("x1", np.uint64), ("x2", np.uint64),
                                                     ("u1", np.float64), ("u2", np.float64),
                                                     ("z", np.float64), ("zs", np.float64))}
    _lib.call("tip_hash_pairs", d_s.ptr, n, bufs["x1"].ptr, bufs["x2"].ptr, bufs["u1"].ptr,
              bufs["u2"].ptr, bufs["z"].ptr, bufs["zs"].ptr, None)
    for k, b in bufs.items():
        out[k] = b.download()
        b.free()
    d_s.free()
    return out


def _hash_rng_pair(seed: int):
    """Scalar mirror of the reference's _hash_rng_pair(seed) -> (u1, u2) (:2322-2336)."""
    r = hash_rng_pairs(np.array([int(seed) & 0xFFFFFFFFFFFFFFFF], np.uint64))
    return float(r["u1"][0]), float(r["u2"][0])


def noise_seed(step: int, seed: int = 0) -> int:
    """Per-step noise seed; seed 0 gives the reference's uint64(step)."""
    return int(_lib.load().tip_noise_seed(int(step) & 0xFFFFFFFFFFFFFFFF,
                                          int(seed) & 0xFFFFFFFFFFFFFFFF))


def vacuum_count(phi, esc: float, phi2=None) -> int:
    """#sites with |phi| > esc (or rho > esc for a complex pair) -- true-vacuum count,
    postprocess/compute_lattice_Tc1.py:24."""
    from .device import Counter

    c = Counter(1)
    arrs = [phi] if phi2 is None else [phi, phi2]
    devs = []
    for a in arrs:
        devs.append(a if _is_dev(a) else DeviceArray.from_numpy(np.ascontiguousarray(a)))
    n = int(np.prod(phi.shape))
    dtype = np.dtype(phi.dtype)
    if phi2 is not None:
        if dtype != np.float32:
            raise TypeError("complex vacuum count takes float32 fields")
        _lib.call("tip_vacuum_count_c32", devs[0].ptr, devs[1].ptr, n, float(esc), c.ptr, None)
    elif dtype == np.float32:
        _lib.call("tip_vacuum_count_f32", devs[0].ptr, n, float(esc), c.ptr, None)
    elif dtype == np.float64:
        _lib.call("tip_vacuum_count_f64", devs[0].ptr, n, float(esc), c.ptr, None)
    else:
        raise TypeError(dtype)
    return int(c.read()[0])


def field_minmax(phi) -> tuple:
    """(min, max) of a field on the device (replaces the per-step phi.min()/phi.max() host
    scans, latticeSimeRescale_numba.py:3655-3656)."""
    d = phi if _is_dev(phi) else DeviceArray.from_numpy(np.ascontiguousarray(phi))
    out = DeviceArray((2,), np.float64)
    name = "tip_minmax_f64" if np.dtype(d.dtype) == np.float64 else "tip_minmax_f32"
    _lib.call(name, d.ptr, int(np.prod(d.shape)), out.ptr, None)
    r = out.download()
    return float(r[0]), float(r[1])
