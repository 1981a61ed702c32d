"""The V' table producer on the device: tip_build_vprime_table.

Reference: Vprime_scalar + build_vprime_table, simulation/latticeSimeRescale_numba.py:805-849
(complex driver simulation/latticeSimComplex_numba.py:818-862).  The host restatement in
tip.potential supplies the model (couplings, _POT_COEFFS) and the dJ splines (their
coefficients are uploaded once per spline); the table itself is evaluated on the GPU in
float64 and stays there.
"""
from __future__ import annotations

import ctypes as C
from typing import Optional, Tuple

import numpy as np

from . import _lib
from . import potential as P
from .device import DeviceArray

_SPLINES: dict = {}  # id(spline) -> (spline, cb on the device, cf on the device)


def model_struct(p: P.ModelParams, sp: P.DJSpline) -> _lib.VprimeModel:
    return _lib.VprimeModel(p.lam, p.mphi, p.mb2, p.yb, p.gb, p.yf, p.gf, p.coeff_b, p.coeff_f,
                            sp.x_min, sp.h, sp.nseg)


def spline_on_device(sp: P.DJSpline) -> Tuple[DeviceArray, DeviceArray]:
    """The spline's (4, nseg) coefficient arrays in HBM (uploaded on first use)."""
    hit = _SPLINES.get(id(sp))
    if hit is None or hit[0] is not sp:
        cb = DeviceArray.from_numpy(np.ascontiguousarray(sp.cb, dtype=np.float64))
        cf = DeviceArray.from_numpy(np.ascontiguousarray(sp.cf, dtype=np.float64))
        hit = (sp, cb, cf)
        _SPLINES[id(sp)] = hit
    return hit[1], hit[2]


def build_vprime_table(T: float, phi_min: float, phi_max: float, n_table: int,
                       p: P.ModelParams, sp: Optional[P.DJSpline] = None,
                       out: Optional[DeviceArray] = None, stream=None
                       ) -> Tuple[DeviceArray, float, float]:
    """(table on the device, tmin, dinv) with table[i] = V'(phi_min + i dphi, T), the device
    twin of tip.potential.build_vprime_table."""
    sp = sp or P.dj_spline()
    cb, cf = spline_on_device(sp)
    if out is None:
        out = DeviceArray((int(n_table),), np.float64)
    elif out.shape != (int(n_table),) or out.dtype != np.float64:
        raise ValueError("out must be a float64 device array of n_table points")
    m = model_struct(p, sp)
    tmin, dinv = C.c_double(0.0), C.c_double(0.0)
    _lib.call("tip_build_vprime_table", out.ptr, int(n_table), float(T), float(phi_min),
              float(phi_max), C.byref(m), cb.ptr, cf.ptr, C.byref(tmin), C.byref(dinv), stream)
    return out, float(tmin.value), float(dinv.value)
