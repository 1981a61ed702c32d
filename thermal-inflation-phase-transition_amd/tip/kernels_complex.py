"""Drop-in replacements for the complex-field kernels of
simulation/latticeSimComplex_numba.py (phi = phi1 + i phi2, U(1) potential in rho with an
optional Z_N term).

  rk2_step_table(phi1, phi2, pi1, pi2, dt, dx, eta_eff, mu, noise1, noise2,
                 phi1_tmp, phi2_tmp, pi1_tmp, pi2_tmp, inv_a2,
                 vp_table, vp_tmin, vp_dinv, vp_npts)                :2067-2215
  overdamped_euler_step_table(phi1, phi2, dt, dx, eta_eff, mu, inv_a2,
                 vp_table, vp_tmin, vp_dinv, vp_npts, noise_scale, seed)  :2422-2517
  baoab_step_table(phi1, phi2, pi1, pi2, dt, dx, gamma, mu, inv_a2,
                 vp_table, vp_tmin, vp_dinv, vp_npts, noise_scale, seed)  :2223-2419
  compute_winding_number(phi1, phi2, winding_out)                    :2683-2744

The Z_N parameters, a module global in the reference (_ZN_PARAMS = [N, delta_V, active],
:227-231), are held in ZN_PARAMS here and may be passed per call.
"""
from __future__ import annotations

import numpy as np

from . import _lib
from .device import Counter, DeviceArray
from .kernels import (BAOAB_CLOSE, BAOAB_OPEN, _check_field, _field_set, _is_dev, _staging,
                      _table_dev, _write_back)

# [N (order), delta_V (strength), active flag] -- same meaning as the reference's global
ZN_PARAMS = np.array([0.0, 0.0, 0.0], dtype=np.float64)


def rk2_step_table(phi1, phi2, pi1, pi2, dt, dx, eta_eff, mu, noise1, noise2, phi1_tmp,
                   phi2_tmp, pi1_tmp, pi2_tmp, inv_a2, vp_table, vp_tmin, vp_dinv, vp_npts, *,
                   zn=None, noise_seed: int = 0, noise_scale: float = 0.0, vac_count=None,
                   vac_esc: float = 1e4, stream=None):
    """One RK2 Langevin step of the complex field, in place on (phi1, phi2, pi1, pi2).

    Reference: rk2_step_table, simulation/latticeSimComplex_numba.py:2067-2215.
    noise1=noise2=None selects inline hash noise (one hash per site, cos/sin components).
    """
    _check_field("phi1", phi1)
    shape, dtype = tuple(phi1.shape), np.dtype(phi1.dtype)
    if dtype not in (np.float32, np.float64):
        raise TypeError(f"phi1: unsupported dtype {dtype}")
    fields = [phi1, phi2, pi1, pi2, phi1_tmp, phi2_tmp, pi1_tmp, pi2_tmp]
    names = ["phi1", "phi2", "pi1", "pi2", "phi1_tmp", "phi2_tmp", "pi1_tmp", "pi2_tmp"]
    for nm, a in zip(names[1:], fields[1:]):
        _check_field(nm, a, shape, dtype)
    if (noise1 is None) != (noise2 is None):
        raise TypeError("noise1 and noise2 must both be arrays or both None")
    if noise1 is not None:
        _check_field("noise1", noise1, shape, dtype)
        _check_field("noise2", noise2, shape, dtype)
    zn = ZN_PARAMS if zn is None else np.asarray(zn, dtype=np.float64)
    table = _table_dev(vp_table, int(vp_npts))
    nx, ny, nz = shape
    host = not _is_dev(phi1)
    if host:
        devs = [_staging.get("c_" + nm, shape, dtype) for nm in names]
        for d, a in zip(devs[:4], fields[:4]):
            d.upload(a)
        n1 = n2 = None
        if noise1 is not None:
            n1 = _staging.get("c_noise1", shape, dtype)
            n2 = _staging.get("c_noise2", shape, dtype)
            n1.upload(noise1)
            n2.upload(noise2)
    else:
        devs = fields
        n1, n2 = noise1, noise2
    name = "tip_rk2_step_table_c64" if dtype == np.float64 else "tip_rk2_step_table_c32"
    _lib.call(name, *(d.ptr for d in devs[:4]), float(dt), float(dx), float(eta_eff), float(mu),
              n1.ptr if n1 is not None else None, n2.ptr if n2 is not None else None,
              *(d.ptr for d in devs[4:]), float(inv_a2), table.ptr, float(vp_tmin),
              float(vp_dinv), int(vp_npts), float(zn[0]), float(zn[1]), float(zn[2]), nx, ny, nz,
              int(noise_seed) & 0xFFFFFFFFFFFFFFFF, float(noise_scale),
              vac_count.ptr if vac_count is not None else None, float(vac_esc), stream)
    if host:
        for d, a in zip(devs[:4], fields[:4]):
            d.download(a)


def _complex_fields(names, arrays):
    _check_field(names[0], arrays[0])
    shape, dtype = tuple(arrays[0].shape), np.dtype(arrays[0].dtype)
    if dtype not in (np.float32, np.float64):
        raise TypeError(f"{names[0]}: unsupported dtype {dtype}")
    for nm, a in zip(names[1:], arrays[1:]):
        _check_field(nm, a, shape, dtype)
    return shape, dtype


def overdamped_euler_step_table(phi1, phi2, dt, dx, eta_eff, mu, inv_a2, vp_table, vp_tmin,
                                vp_dinv, vp_npts, noise_scale, seed, *, zn=None, stream=None):
    """One overdamped Euler-Maruyama step of the complex field, in place on (phi1, phi2).

    Reference: overdamped_euler_step_table, simulation/latticeSimComplex_numba.py:2422-2517,
    evaluated on one consistent field; noise (cos, sin) of the hash draw of noise_seed = seed.
    """
    shape, dtype = _complex_fields(["phi1", "phi2"], [phi1, phi2])
    zn = ZN_PARAMS if zn is None else np.asarray(zn, dtype=np.float64)
    table = _table_dev(vp_table, int(vp_npts))
    devs, host = _field_set(["cod_phi1", "cod_phi2"], [phi1, phi2], dtype)
    outs = [_staging.get(nm, shape, dtype) for nm in ("cod_out1", "cod_out2")]
    name = "tip_overdamped_c64" if dtype == np.float64 else "tip_overdamped_c32"
    _lib.call(name, devs[0].ptr, devs[1].ptr, outs[0].ptr, outs[1].ptr, float(dt), float(dx),
              float(eta_eff), float(mu), float(inv_a2), table.ptr, float(vp_tmin),
              float(vp_dinv), int(vp_npts), float(zn[0]), float(zn[1]), float(zn[2]), *shape,
              int(seed) & 0xFFFFFFFFFFFFFFFF, float(noise_scale), stream)
    _write_back(host, [phi1, phi2], outs)


def baoab_step_table(phi1, phi2, pi1, pi2, dt, dx, gamma, mu, inv_a2, vp_table, vp_tmin,
                     vp_dinv, vp_npts, noise_scale, seed, *, zn=None, stream=None):
    """One BAOAB step of the complex field, in place on (phi1, phi2, pi1, pi2).

    Reference: baoab_step_table, simulation/latticeSimComplex_numba.py:2223-2419, with every
    force evaluated on one consistent field; noise_scale is the reference's sqrt(T / (mu^2
    dx^3)) (:4313).
    """
    names = ["cba_phi1", "cba_phi2", "cba_pi1", "cba_pi2"]
    fields = [phi1, phi2, pi1, pi2]
    shape, dtype = _complex_fields(["phi1", "phi2", "pi1", "pi2"], fields)
    zn = ZN_PARAMS if zn is None else np.asarray(zn, dtype=np.float64)
    table = _table_dev(vp_table, int(vp_npts))
    devs, host = _field_set(names, fields, dtype)
    outs = [_staging.get(nm + "_out", shape, dtype) for nm in names]
    name = "tip_baoab_c64" if dtype == np.float64 else "tip_baoab_c32"
    tail = (float(dt), float(dx), float(gamma), float(mu), float(inv_a2), float(inv_a2),
            table.ptr, float(vp_tmin), float(vp_dinv), int(vp_npts), float(zn[0]),
            float(zn[1]), float(zn[2]), *shape)
    _lib.call(name, BAOAB_OPEN, *(d.ptr for d in devs), *(o.ptr for o in outs), *tail,
              int(seed) & 0xFFFFFFFFFFFFFFFF, float(noise_scale), stream)
    _lib.call(name, BAOAB_CLOSE, outs[0].ptr, outs[1].ptr, outs[2].ptr, outs[3].ptr, None, None,
              outs[2].ptr, outs[3].ptr, *tail, 0, 0.0, stream)
    _write_back(host, fields, outs)


def compute_winding_number(phi1, phi2, winding_out=None, *, plaquettes: bool = False):
    """Per-site winding number density (sum over the XY, XZ, YZ plaquettes).

    Reference: compute_winding_number, simulation/latticeSimComplex_numba.py:2683-2744;
    string count int(sum(|W| > 0.5)) of :4661.  Fills winding_out (float32, as the
    reference's winding_buf) if given and returns (winding, n_strings, n_plaquettes, plaq)
    where plaq is the int8 [3, nx, ny, nz] per-plaquette winding (XY, XZ, YZ) when
    plaquettes=True, else None.
    """
    _check_field("phi1", phi1)
    shape, dtype = tuple(phi1.shape), np.dtype(phi1.dtype)
    _check_field("phi2", phi2, shape, dtype)
    nx, ny, nz = shape
    d1 = phi1 if _is_dev(phi1) else DeviceArray.from_numpy(phi1)
    d2 = phi2 if _is_dev(phi2) else DeviceArray.from_numpy(phi2)
    w = DeviceArray(shape, np.float32)
    plaq = DeviceArray((3,) + shape, np.int8) if plaquettes else None
    cnt = Counter(2)
    name = "tip_winding_f64" if dtype == np.float64 else "tip_winding_f32"
    if dtype not in (np.float32, np.float64):
        raise TypeError(dtype)
    _lib.call(name, d1.ptr, d2.ptr, plaq.ptr if plaq is not None else None, w.ptr, cnt.ptr, nx,
              ny, nz, None)
    wh = w.download() if winding_out is None else w.download(winding_out)
    counts = cnt.read()
    ph = plaq.download() if plaq is not None else None
    return wh, int(counts[0]), int(counts[1]), ph
