"""Finite-temperature potential and the V'(phi) table producer (host, float64).

Restates the reference's table producer:
  dJ spline        simulation/latticeSimeRescale_numba.py:432-455
                   (256 uniform points u in [0, 100], not-a-knot cubic spline of dJ_b, dJ_f)
  Vprime_scalar    :805-837     build_vprime_table :840-849
  _POT_COEFFS      real :231-241 ([1,1] V_correct x nb,nf);  complex
                   simulation/latticeSimComplex_numba.py:215-221 ([2,-1] V_p, no nb/nf)
  thermal VEV      :852-863 (brentq);  escape point :869-943
  noise amplitude  :3714-3718;  Hubble schedule :3619-3647, :706-717

The reference gets dJ_b(u), dJ_f(u) by quadrature in cosmoTransitions.finiteT
(dJb_exact/dJf_exact, version unpinned, not installed here).  This module evaluates the
same functions from their Bessel series

  J_b(u)   = -sum_n u^2 K_2(n u) / n^2        dJ_b/du = sum_n u^2 K_1(n u) / n
  J_f(u)   = -sum_n (-1)^(n+1) u^2 K_2(n u)/n^2   dJ_f/du = sum_n (-1)^(n+1) u^2 K_1(n u) / n

(the convention in which both species raise the thermal mass), which the tests pin against
direct quadrature and against V' values printed in the reference's run logs.

The full FP64 potential with the Coleman-Weinberg term (cosmolattice_ext/models/
thermal_tables.hpp:96-146) is `vprime_full`, the ground truth of the learned surrogate.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field, replace
from typing import Optional, Tuple

import numpy as np
from scipy.interpolate import CubicSpline
from scipy.optimize import brentq
from scipy.special import k1, k1e, kve

YMAX = 100.0  # latticeSimeRescale_numba.py:432
N_Y = 256     # :433


# ------------------------------------------------------------------------------------------
# Thermal functions
# ------------------------------------------------------------------------------------------
def _bessel_sum(u: np.ndarray, order: int, fermion: bool, power: int) -> np.ndarray:
    """sum_{n>=1} s_n * u^2 * K_order(n u) / n^power, s_n = 1 (boson) or (-1)^(n+1).

    Summed with exponentially scaled Bessel functions until the terms fall below 1e-18 of
    the running sum; u = 0 is the exact limit (0 for dJ, -pi^4/45 or -7pi^4/360 for J).
    """
    u = np.atleast_1d(np.asarray(u, dtype=np.float64))
    out = np.zeros_like(u)
    pos = u > 0
    x = u[pos]
    acc = np.zeros_like(x)
    n = 1
    active = np.ones_like(x, dtype=bool)
    while active.any() and n < 200000:
        nx = n * x[active]
        if order == 1:
            kv = k1e(nx) * np.exp(-nx)
        else:
            kv = kve(order, nx) * np.exp(-nx)
        term = x[active] ** 2 * kv / n ** power
        if fermion and n % 2 == 0:
            term = -term
        acc[active] += term
        small = np.abs(term) <= 1e-18 * np.maximum(np.abs(acc[active]), 1e-300)
        idx = np.flatnonzero(active)
        active[idx[small]] = False
        n += 1
    out[pos] = acc
    return out


def dJb(u) -> np.ndarray:
    """dJ_b/du (cosmoTransitions dJb_exact convention, u = m/T)."""
    return _bessel_sum(u, 1, False, 1)


def dJf(u) -> np.ndarray:
    """dJ_f/du (cosmoTransitions dJf_exact convention, u = m/T)."""
    return _bessel_sum(u, 1, True, 1)


def d2J(u, fermion: bool) -> np.ndarray:
    """d^2 J/du^2 = dJ/du / u - sum_n s_n u^2 K_0(n u) (from d/dx[x K_1(x)] = -x K_0(x));
    pi^2/6 (boson) or pi^2/12 (fermion) at u = 0."""
    u = np.atleast_1d(np.asarray(u, dtype=np.float64))
    out = np.empty_like(u)
    pos = u > 0
    d1 = dJf(u[pos]) if fermion else dJb(u[pos])
    out[pos] = d1 / u[pos] - _bessel_sum(u[pos], 0, fermion, 0)
    out[~pos] = math.pi ** 2 / (12.0 if fermion else 6.0)
    return out


def Jb(u) -> np.ndarray:
    u = np.atleast_1d(np.asarray(u, dtype=np.float64))
    out = -_bessel_sum(u, 2, False, 2)
    out[u == 0] = -math.pi ** 4 / 45.0
    return out


def Jf(u) -> np.ndarray:
    u = np.atleast_1d(np.asarray(u, dtype=np.float64))
    out = -_bessel_sum(u, 2, True, 2)
    out[u == 0] = -7.0 * math.pi ** 4 / 360.0
    return out


@dataclass(frozen=True)
class DJSpline:
    """Piecewise-cubic coefficients of dJ_b and dJ_f on the uniform u grid
    (latticeSimeRescale_numba.py:432-455: cs.c rows = powers [3, 2, 1, 0])."""

    cb: np.ndarray  # (4, nseg)
    cf: np.ndarray  # (4, nseg)
    x_min: float
    h: float
    nseg: int

    @property
    def xhi_clamp(self) -> float:
        return self.x_min + self.h * self.nseg - 1e-12


_SPLINE_CACHE: dict = {}


def dJ_soft(u, x_cut: float, fermion: bool) -> np.ndarray:
    """Soft-mode (k < pi/dx) part of dJ/du as the reference's lattice counterterm subtracts it
    (latticeSimeRescale_numba.py:460-490): (1/2pi^2) int_0^x_cut x^2 u / (z (e^z -+ 1)) dx,
    z = sqrt(x^2 + u^2), integrand cut to 0 for z > 50; 0 for x_cut < 1e-10 or u < 1e-30."""
    from scipy.integrate import quad
    u = np.atleast_1d(np.asarray(u, dtype=np.float64))
    out = np.zeros_like(u)
    if x_cut < 1e-10:
        return out
    sgn = 1.0 if fermion else -1.0
    for i, y in enumerate(u):
        if y < 1e-30:
            continue

        def f(x, y=y):
            z = math.sqrt(x * x + y * y)
            return 0.0 if z > 50.0 else x * x * y / (z * (math.exp(z) + sgn))

        out[i] = quad(f, 0.0, x_cut, limit=200)[0] / (2.0 * math.pi ** 2)
    return out


def counterterm_xcut(dx_phys: float, T0: float) -> float:
    """x_cut = k_max / T0 with k_max = pi / dx_phys (latticeSimeRescale_numba.py:759)."""
    return math.pi / (dx_phys * T0)


def dj_spline(n_y: int = N_Y, ymax: float = YMAX, x_cut: float = 0.0) -> DJSpline:
    """The dJ_b, dJ_f splines; x_cut > 0 applies the lattice counterterm (the soft-mode part
    subtracted from the grid values before the spline, :758-795)."""
    key = (n_y, ymax, x_cut)
    if key not in _SPLINE_CACHE:
        grid = np.linspace(0.0, ymax, n_y, dtype=np.float64)
        gb, gf = dJb(grid), dJf(grid)
        if x_cut > 0.0:
            gb = gb - dJ_soft(grid, x_cut, fermion=False)
            gf = gf - dJ_soft(grid, x_cut, fermion=True)
        cs_b = CubicSpline(grid, gb, bc_type="not-a-knot")
        cs_f = CubicSpline(grid, gf, bc_type="not-a-knot")
        _SPLINE_CACHE[key] = DJSpline(
            cb=np.ascontiguousarray(cs_b.c, dtype=np.float64),
            cf=np.ascontiguousarray(cs_f.c, dtype=np.float64),
            x_min=float(grid[0]), h=float(grid[1] - grid[0]), nseg=int(n_y - 1))
    return _SPLINE_CACHE[key]


def _spline_eval(c: np.ndarray, sp: DJSpline, x: np.ndarray) -> np.ndarray:
    """Vprime_scalar's clamp + Horner (latticeSimeRescale_numba.py:815-820)."""
    xc = np.minimum(np.maximum(x, sp.x_min), sp.x_min + sp.h * sp.nseg - 1e-12)
    idx = ((xc - sp.x_min) / sp.h).astype(np.int64)
    idx = np.clip(idx, 0, sp.nseg - 1)
    ds = xc - (sp.x_min + idx * sp.h)
    return ((c[0, idx] * ds + c[1, idx]) * ds + c[2, idx]) * ds + c[3, idx]


# ------------------------------------------------------------------------------------------
# Model parameters
# ------------------------------------------------------------------------------------------
@dataclass(frozen=True)
class ModelParams:
    """Couplings of the reference drivers (latticeSimeRescale_numba.py:642-643,720-724)."""

    lam: float = 1e-24
    mphi: float = 1000.0
    mb2: float = 1.0e6           # bosonMassSquared
    yb: float = 1.09             # bosonCoupling
    gb: float = 1.05             # bosonGaugeCoupling
    yf: float = 1.09             # fermionCoupling
    gf: float = 1.05             # fermionGaugeCoupling
    coeff_b: float = 20.0        # _POT_COEFFS[0] (convention coefficient x nb)
    coeff_f: float = 20.0        # _POT_COEFFS[1]

    @staticmethod
    def real(potential_type: str = "V_correct", nb: float = 20.0, nf: float = 20.0,
             lam: float = 1e-24, **kw) -> "ModelParams":
        """Real driver: _POT_COEFFS = convention x (nb, nf), :231-241 (defaults nb=nf=20)."""
        cb, cf = {"V_correct": (1.0, 1.0), "V_p": (2.0, -1.0),
                  "fermion_only": (0.0, 1.0)}[potential_type]
        return ModelParams(lam=lam, coeff_b=cb * nb, coeff_f=cf * nf, **kw)

    @staticmethod
    def complex(potential_type: str = "V_p", lam: float = 1e-16, **kw) -> "ModelParams":
        """Complex driver: _POT_COEFFS default [2,-1] (V_p), no multiplicities,
        latticeSimComplex_numba.py:215-221; lam = 1e-16 (:638)."""
        cb, cf = {"V_correct": (1.0, 1.0), "V_p": (2.0, -1.0),
                  "fermion_only": (0.0, 1.0)}[potential_type]
        return ModelParams(lam=lam, coeff_b=cb, coeff_f=cf, **kw)

    @property
    def vev_tree(self) -> float:
        return math.sqrt(self.mphi ** 2 / self.lam)


def vprime(phi, T: float, p: ModelParams, sp: Optional[DJSpline] = None) -> np.ndarray:
    """Vectorised Vprime_scalar(phi, T) (latticeSimeRescale_numba.py:805-837): tree +
    thermal boson + thermal fermion through the dJ spline; no Coleman-Weinberg term."""
    sp = sp or dj_spline()
    phi = np.asarray(phi, dtype=np.float64)
    dV = p.lam * phi ** 3 - p.mphi ** 2 * phi
    pref = T ** 4 / (2.0 * np.pi ** 2)
    gb2, gg2 = p.yb ** 2, p.gb ** 2
    coef_b = 0.25 * gb2 + (2.0 / 3.0) * gg2
    xb_sq = p.mb2 + 0.5 * gb2 * phi ** 2 + coef_b * T ** 2
    with np.errstate(invalid="ignore", divide="ignore"):
        xb = np.sqrt(np.where(xb_sq > 0, xb_sq, 0.0)) / T
        dJb_v = _spline_eval(sp.cb, sp, xb)
        dxb = 0.5 * gb2 * phi / (T ** 2 * np.maximum(xb, 1e-20))
        dV = dV + np.where(xb_sq > 0, pref * p.coeff_b * dJb_v * dxb, 0.0)
        gf2, gfg2 = p.yf ** 2, p.gf ** 2
        xf_sq = 0.5 * gf2 * phi ** 2 + (1.0 / 6.0) * gfg2 * T ** 2
        xf = np.sqrt(np.where(xf_sq > 0, xf_sq, 0.0)) / T
        dJf_v = _spline_eval(sp.cf, sp, xf)
        dxf = 0.5 * gf2 * phi / (T ** 2 * np.maximum(xf, 1e-20))
        dV = dV + np.where(xf_sq > 0, pref * p.coeff_f * dJf_v * dxf, 0.0)
    return dV


def vprime_scalar(phi: float, T: float, p: ModelParams, sp: Optional[DJSpline] = None) -> float:
    return float(vprime(np.array([phi]), T, p, sp)[0])


def build_vprime_table(T: float, phi_min: float, phi_max: float, n_table: int,
                       p: ModelParams, sp: Optional[DJSpline] = None
                       ) -> Tuple[np.ndarray, float, float]:
    """(table, tmin, dinv) with table[i] = V'(phi_min + i*dphi) (latticeSimeRescale_numba.py:840-849)."""
    grid = np.linspace(phi_min, phi_max, n_table, dtype=np.float64)
    table = vprime(grid, T, p, sp)
    return table, float(phi_min), 1.0 / ((phi_max - phi_min) / (n_table - 1))


def thermal_vev(T: float, p: ModelParams, sp: Optional[DJSpline] = None) -> float:
    """brentq root of V' on [1e-3 vev, 2 vev] (latticeSimeRescale_numba.py:852-863)."""
    v = p.vev_tree
    try:
        return brentq(lambda x: vprime_scalar(x, T, p, sp), v * 0.001, v * 2.0)
    except ValueError:
        return v


def barrier_top(T: float, p: ModelParams, vev_th: Optional[float] = None,
                sp: Optional[DJSpline] = None) -> Optional[float]:
    """First + -> - sign change of V' on the log grid of latticeSimeRescale_numba.py:877-917."""
    vev_th = vev_th or thermal_vev(T, p, sp)
    pts = np.concatenate([[1e-10], np.logspace(0, np.log10(vev_th * 0.99), 500)])
    vals = vprime(pts, T, p, sp)
    for i in range(len(pts) - 1):
        if vals[i] > 0 and vals[i + 1] < 0:
            return brentq(lambda x: vprime_scalar(x, T, p, sp), pts[i], pts[i + 1])
    return None


# ------------------------------------------------------------------------------------------
# Full FP64 potential with the Coleman-Weinberg term (surrogate ground truth)
# cosmolattice_ext/models/thermal_tables.hpp:96-146, 212-224: dV_CW = dm2*m2/(64 pi^2) *
# 2 (ln(|m2|/T) - 1), m2 = 3 lam phi^2, dm2 = 6 lam phi; thermal part with exact dJ.
# ------------------------------------------------------------------------------------------
def vprime_full(phi, T, p: ModelParams, include_cw: bool = True) -> np.ndarray:
    phi = np.asarray(phi, dtype=np.float64)
    T = np.asarray(T, dtype=np.float64)
    pref = T ** 4 / (2.0 * np.pi ** 2)
    ub = np.sqrt(np.maximum(p.mb2 + 0.5 * p.yb ** 2 * phi ** 2
                            + (0.25 * p.yb ** 2 + 2.0 / 3.0 * p.gb ** 2) * T ** 2, 0.0)) / T
    uf = np.sqrt(np.maximum(0.5 * p.yf ** 2 * phi ** 2 + (1.0 / 6.0) * p.gf ** 2 * T ** 2, 0.0)) / T
    dub = 0.5 * p.yb ** 2 * phi / (T * T * np.maximum(ub, 1e-20))
    duf = 0.5 * p.yf ** 2 * phi / (T * T * np.maximum(uf, 1e-20))
    vp = p.lam * phi ** 3 - p.mphi ** 2 * phi
    shape = np.broadcast(phi, T).shape
    vp = vp + pref * (p.coeff_b * dJb(ub.ravel()).reshape(shape) * dub
                      + p.coeff_f * dJf(uf.ravel()).reshape(shape) * duf)
    if include_cw:
        m2 = 3.0 * p.lam * phi ** 2
        dm2 = 6.0 * p.lam * phi
        with np.errstate(divide="ignore", invalid="ignore"):
            cw = np.where(np.abs(m2) > 0,
                          dm2 * m2 / (64.0 * np.pi ** 2) * 2.0 * (np.log(np.abs(m2) / T) - 1.0), 0.0)
        vp = vp + cw
    return vp


# ------------------------------------------------------------------------------------------
# Schedules and amplitudes
# ------------------------------------------------------------------------------------------
G_STAR = 106.75
M_PL = 2.4e18
DEL_V_REAL = 1e36 / 4   # latticeSimeRescale_numba.py:427


def noise_scale(eta_eff: float, T: float, dt: float, mu: float, dx_phys: float,
                inv_a3: float = 1.0) -> float:
    """sqrt(2 eta_eff T dt inv_a3 / (mu^2 dx_phys^3)), latticeSimeRescale_numba.py:3714-3718."""
    return math.sqrt(2.0 * eta_eff * T * dt * inv_a3 / (mu ** 2 * dx_phys ** 3))


@dataclass
class HubbleState:
    """Per-step T / H / scale-factor recurrence, latticeSimeRescale_numba.py:3622-3637."""

    T0: float
    mu: float
    eta: float
    del_v: float = DEL_V_REAL
    a: float = 1.0

    def step(self, dt: float):
        T = self.T0 / self.a
        H = math.sqrt((T ** 4 * (G_STAR * math.pi ** 2 / 30.0) + self.del_v) / (3.0 * M_PL ** 2))
        eta_eff = self.eta + 3.0 * H / self.mu
        inv_a2 = 1.0 / (self.a * self.a)
        inv_a3 = 1.0 / (self.a * self.a * self.a)
        self.a += self.a * H * (dt / self.mu)
        T_mid = self.T0 / self.a
        return dict(T=T, H=H, eta_eff=eta_eff, inv_a2=inv_a2, inv_a3=inv_a3, T_mid=T_mid)
