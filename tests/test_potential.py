"""The V'(phi, T) producer (host float64) against what pins it.

  * J_b, J_f and their derivatives from Bessel series against direct quadrature of
    int y^2 ln(1 -+ exp(-sqrt(y^2 + u^2))) dy, and the u -> 0 limits
  * V' values, barrier tops and counterterm soft-mode fractions printed by the reference's
    own runs (tests/golden/log_vprime_pins.json, 29 distinct cases extracted from its logs):
    every printed digit reproduced -- this also fixes the cosmoTransitions sign convention
    and the lattice counterterm
  * the table-path spline: exact at its knots, its interpolation error bounded
  * the full potential with the Coleman-Weinberg term (the surrogate's truth) against the
    reference's own C++ assembly (cosmolattice_ext/models/thermal_tables.hpp through its
    test programs, built into oracle/_ref when the checkout is present)
"""
import json
import math
import re
import struct
import subprocess
from pathlib import Path

import numpy as np
import pytest
from scipy.integrate import quad

from tip import potential as P

ROOT = Path(__file__).resolve().parents[1]
PINS = json.loads((ROOT / "tests" / "golden" / "log_vprime_pins.json").read_text())
REF = ROOT / "oracle" / "_ref"


def _J_quad(u, fermion):
    s = 1.0 if fermion else -1.0
    f = lambda y: y * y * math.log1p(s * math.exp(-math.sqrt(y * y + u * u)))
    v = quad(f, 0.0, np.inf, limit=400, epsabs=0, epsrel=1e-13)[0]
    return -v if fermion else v  # J_f carries the minus sign (thermal mass > 0)


def _dJ_quad(u, fermion):
    s = 1.0 if fermion else -1.0
    def f(y):
        z = math.sqrt(y * y + u * u)
        e = math.exp(-z)  # y^2 u / (z (e^z -+ 1)) written without overflow
        return y * y * u * e / (z * (1.0 + s * e))
    return quad(f, 0.0, np.inf, limit=400, epsabs=0, epsrel=1e-13)[0]


@pytest.mark.parametrize("u", [0.05, 0.5, 1.0, 2.7, 10.0, 30.0])
def test_series_match_quadrature(u):
    for fermion, J, dJ in ((False, P.Jb, P.dJb), (True, P.Jf, P.dJf)):
        jq, dq = _J_quad(u, fermion), _dJ_quad(u, fermion)
        assert abs(J(u)[0] - jq) <= 1e-10 * max(abs(jq), 1e-30), (u, fermion)
        assert abs(dJ(u)[0] - dq) <= 1e-10 * max(abs(dq), 1e-30), (u, fermion)
        # second derivative against a central difference of the first
        h = 1e-4 * max(u, 1.0)
        fd = (dJ(u + h)[0] - dJ(u - h)[0]) / (2 * h)
        assert abs(P.d2J(u, fermion)[0] - fd) <= 1e-5 * max(abs(fd), 1e-12)


def test_small_u_limits():
    assert P.Jb(0.0)[0] == -math.pi ** 4 / 45.0
    assert P.Jf(0.0)[0] == -7.0 * math.pi ** 4 / 360.0
    assert P.dJb(0.0)[0] == 0.0 and P.dJf(0.0)[0] == 0.0
    u = 1e-3  # dJ_b ~ pi^2 u / 6 - pi u^2 / 2, dJ_f ~ pi^2 u / 12
    assert abs(P.dJb(u)[0] - (math.pi ** 2 * u / 6 - math.pi * u * u / 2)) < 1e-5 * u  # O(u^3 ln u)
    assert abs(P.dJf(u)[0] - math.pi ** 2 * u / 12) < 1e-5 * u


def _model(rec, ptype):
    return P.ModelParams.complex(ptype, lam=1e6 / rec["vev_tree"] ** 2, yb=rec["yb"],
                                 yf=rec["yf"])


def _spline(rec):
    ct = rec["counterterm"]
    return P.dj_spline(x_cut=P.counterterm_xcut(ct["dx_phys"], ct["T0"])) if ct else None


def _printed_ulp(v):
    """Half a unit in the last of the 5 printed significant figures (%.4e)."""
    return 0.5 * 10.0 ** (math.floor(math.log10(abs(v))) - 4)


@pytest.mark.parametrize("rec", PINS, ids=[r["log"] for r in PINS])
def test_log_printed_vprime(rec):
    """Logs without a potential-type line predate the flag and ran V_p ([2, -1]); only that
    convention reproduces them."""
    ptype = rec["potential_type"] or "V_p"
    m, sp = _model(rec, ptype), _spline(rec)
    for phi, v in rec["vprime"].items():
        ours = P.vprime_scalar(float(phi), rec["T"], m, sp)
        assert abs(ours - v) <= 1.05 * _printed_ulp(v), (phi, ours, v)
    if rec["barrier_top"] is not None:
        bt = P.barrier_top(rec["T"], m, sp=sp)
        assert abs(bt - rec["barrier_top"]) <= 1.05 * _printed_ulp(rec["barrier_top"])
    if rec["potential_type"] is None:
        for other in ("V_correct", "fermion_only"):
            o = P.vprime_scalar(1.0, rec["T"], _model(rec, other), sp)
            assert abs(o - rec["vprime"]["1e0"]) > 100 * _printed_ulp(rec["vprime"]["1e0"])


def test_counterterm_soft_fractions():
    """Printed soft-mode fractions sum|dJ_soft| / sum|dJ| on the 256-point grid (4 decimals)."""
    grid = np.linspace(0.0, P.YMAX, P.N_Y)
    seen = set()
    for rec in PINS:
        ct = rec["counterterm"]
        if not ct or ct["frac_b"] is None or (ct["dx_phys"], ct["T0"]) in seen:
            continue
        seen.add((ct["dx_phys"], ct["T0"]))
        x = P.counterterm_xcut(ct["dx_phys"], ct["T0"])
        fb = np.abs(P.dJ_soft(grid, x, False)).sum() / np.abs(P.dJb(grid)).sum()
        ff = np.abs(P.dJ_soft(grid, x, True)).sum() / np.abs(P.dJf(grid)).sum()
        assert abs(fb - ct["frac_b"]) <= 0.5e-4 + 1e-9 and abs(ff - ct["frac_f"]) <= 0.5e-4 + 1e-9
    assert len(seen) >= 3


def test_table_spline_knots_and_interpolation_error():
    """The 256-point not-a-knot spline of the table path (latticeSimeRescale_numba.py:432-455)
    passes through the series at its knots; between knots (h = 100/255) its error is that of a
    cubic through a function with u^3 ln u terms at small u -- up to ~3e-3 absolute below u = 1
    (a few per cent of dJ there), which is how far the reference's table path sits from the
    exact thermal functions the surrogate is trained on."""
    sp = P.dj_spline()
    g = np.linspace(0.0, P.YMAX, P.N_Y)[:-1]
    for c, f in ((sp.cb, P.dJb), (sp.cf, P.dJf)):
        assert np.max(np.abs(P._spline_eval(c, sp, g) - f(g))) <= 1e-15 * np.abs(f(g)).max()
        for lo, hi, bound in ((0.0, 1.0, 3e-3), (1.0, 5.0, 2e-4), (5.0, 100.0, 1e-6)):
            u = np.linspace(lo, hi, 4001)[1:-1]
            assert np.max(np.abs(P._spline_eval(c, sp, u) - f(u))) < bound


def _write_bin(path, umax=150.0, npts=32768):
    """The binary J table in the layout cosmolattice_ext/models/thermal_tables.hpp:66-86 reads
    (int64 n, f64 umin, f64 umax, then Jb, Jf, dJb, dJf, d2Jb, d2Jf), from the series."""
    u = np.linspace(0.0, umax, npts)
    arrs = (P.Jb(u), P.Jf(u), P.dJb(u), P.dJf(u), P.d2J(u, False), P.d2J(u, True))
    with open(path, "wb") as fh:
        fh.write(struct.pack("<q", npts) + struct.pack("<d", u[0]) + struct.pack("<d", u[-1]))
        for a in arrs:
            fh.write(np.asarray(a, dtype="<f8").tobytes())


def _run(exe, *args):
    out = subprocess.run([str(exe), *map(str, args)], capture_output=True, text=True, check=True)
    return out.stdout


def _fields(line):
    return {k: float(v) for k, v in re.findall(r"(\w+)=\s*([-+0-9.eE]+)", line)}


def _tol(vp, phi, m):
    """1e-7 relative, plus 1e-9 of the tree scale |phi| m^2 where tree and thermal terms
    cancel.  The table handed to the reference is 32768 points on [0, 150]: its cubic Hermite
    interpolation error scales as du^4 (1.5e-4 worst relative at the 4096 points of the
    reference's default grid, 3e-8 at 32768), so what remains is the assembly itself."""
    return 1e-7 * abs(vp) + 1e-9 * abs(phi) * m.mphi ** 2


def _ref_model(include_cw=True):
    mpl, gamma = 2.4e18, 4.1667e-4  # the programs' parameter set B
    return P.ModelParams(lam=1e6 / (gamma * mpl) ** 2, coeff_b=20.0, coeff_f=20.0)


@pytest.mark.skipif(not (REF / "vprime_audit").exists(), reason="oracle/_ref not built")
def test_full_potential_vs_reference_cpp(tmp_path):
    binp = tmp_path / "thermal_tables.bin"
    _write_bin(binp)
    m = _ref_model()
    checked = 0
    for cw in (0, 1):
        T = None
        for line in _run(REF / "vprime_audit", binp, cw).splitlines():
            f = _fields(line)
            if "T" in f:
                T = f["T"]
                continue
            phi, vp = f["phi"], f["Vp"]
            ours = float(P.vprime_full(phi, T, m, include_cw=bool(cw)))
            assert abs(ours - vp) <= _tol(vp, phi, m), (cw, T, phi, ours, vp)
            checked += 1
    for line in _run(REF / "test_thermal_tables", binp).splitlines():
        f = _fields(line)
        T, phi, vp = f["T"], f["phi"], f["Vp"]
        ours = float(P.vprime_full(phi, T, m, include_cw=True))
        assert abs(ours - vp) <= _tol(vp, phi, m), (T, phi, ours, vp)
        checked += 1
    assert checked == 2 * 2 * 19 + 6
