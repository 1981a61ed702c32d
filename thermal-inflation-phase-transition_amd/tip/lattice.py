"""Device-resident lattice (the performance path of the hot loop).

A Lattice owns ping-pong (phi, pi) buffers in HBM (complex: phi1, phi2, pi1, pi2), the V'
table in float64 and float32-segment form, and a HIP stream.  step(n) runs n fused RK2
kernels back to back with the reference's hash noise (noise_seed = step ^ seed*golden, the
reference's uint64(step) for seed 0, simulation/latticeSimeRescale_numba.py:3738), with no
host round trip; observables (true-vacuum count, winding, min/max) run on the device.
"""
from __future__ import annotations

import ctypes as C
from typing import Optional

import numpy as np

from . import _lib
from ._lib import StepParams

REAL, COMPLEX = 0, 1


KERNEL_AUTO, KERNEL_ROWS, KERNEL_FLAT = 0, 1, 2


class Lattice:
    def __init__(self, shape, dtype=np.float32, complex_field: bool = False):
        self.shape = tuple(int(s) for s in shape)
        self.dtype = np.dtype(dtype)
        self.complex = bool(complex_field)
        h = C.c_void_p()
        _lib.call("tip_lattice_create", C.byref(h), COMPLEX if self.complex else REAL,
                  self.dtype.itemsize, *self.shape)
        self._h = h
        self.step_count = 0
        self.table_info = None

    # -- state ---------------------------------------------------------------------------
    @property
    def handle(self):
        return self._h

    def _idx(self, name: str) -> int:
        names = ["phi1", "pi1", "phi2", "pi2"] if self.complex else ["phi", "pi"]
        return names.index(name)

    def upload(self, name: str, x: np.ndarray) -> None:
        x = np.ascontiguousarray(x, dtype=self.dtype)
        if x.shape != self.shape:
            raise ValueError(f"{name}: shape {x.shape} != {self.shape}")
        _lib.call("tip_lattice_upload", self._h, self._idx(name), x.ctypes.data)

    def download(self, name: str) -> np.ndarray:
        out = np.empty(self.shape, self.dtype)
        _lib.call("tip_lattice_download", self._h, self._idx(name), out.ctypes.data)
        return out

    def field_ptr(self, name: str) -> int:
        p = C.c_void_p()
        _lib.call("tip_lattice_field_ptr", self._h, self._idx(name), C.byref(p))
        return int(p.value)

    def set_table(self, table: np.ndarray, tmin: float, dinv: float, odd: bool = False) -> None:
        t = np.ascontiguousarray(table, dtype=np.float64)
        _lib.call("tip_lattice_set_table", self._h, t.ctypes.data, t.size, float(tmin),
                  float(dinv), int(bool(odd)))
        self.table_info = (float(t[0]), float(tmin), float(dinv), t.size, bool(odd))

    def build_table(self, T: float, phi_min: float, phi_max: float, npts: int, model,
                    spline=None, odd: bool = False) -> tuple:
        """Build the V' table of (T, model) on the device, straight into the lattice's table
        buffer (tip_lattice_build_table; no host evaluation, no upload).  Returns (tmin,
        dinv) as the reference's build_vprime_table does."""
        from . import potential as P
        from . import vprime_device as VD
        sp = spline or P.dj_spline()
        cb, cf = VD.spline_on_device(sp)
        m = VD.model_struct(model, sp)
        _lib.call("tip_lattice_build_table", self._h, int(npts), float(T), float(phi_min),
                  float(phi_max), C.byref(m), cb.ptr, cf.ptr, int(bool(odd)))
        tmin, dinv = float(phi_min), 1.0 / ((float(phi_max) - float(phi_min)) / (int(npts) - 1))
        self.table_info = (None, tmin, dinv, int(npts), bool(odd))
        return tmin, dinv

    def set_zn(self, n: float, dv: float, active: float) -> None:
        _lib.call("tip_lattice_set_zn", self._h, float(n), float(dv), float(active))

    def set_geometry(self, rows: int = 0, planes_per_block: int = 0, xcd_remap: bool = True,
                     kernel: int = KERNEL_AUTO) -> None:
        """Launch geometry of the float32 step kernel (0 = default).  kernel KERNEL_ROWS: `rows`
        = y rows per thread of the LDS row-tile kernel (real field); KERNEL_FLAT: `rows` = rows
        of 64 z-sites per block of the flat-tile kernel; KERNEL_AUTO: row tiles where they
        apply."""
        _lib.call("tip_lattice_set_geometry", self._h, int(kernel), int(rows),
                  int(planes_per_block), int(bool(xcd_remap)))

    # -- evolution -----------------------------------------------------------------------
    def step(self, nsteps: int, *, dt: float, dx: float, eta_eff: float, mu: float,
             inv_a2: float = 1.0, noise_scale: float = 0.0, seed: int = 0,
             step0: Optional[int] = None, timed: bool = False) -> Optional[float]:
        """Advance nsteps; returns the HIP-event milliseconds of the launches if timed."""
        p = StepParams(dt, dx, eta_eff, mu, inv_a2, noise_scale)
        s0 = self.step_count if step0 is None else int(step0)
        ms = C.c_float(0.0)
        _lib.call("tip_lattice_step", self._h, C.byref(p), int(nsteps), s0 & 0xFFFFFFFFFFFFFFFF,
                  int(seed) & 0xFFFFFFFFFFFFFFFF, C.byref(ms) if timed else None)
        self.step_count = s0 + int(nsteps)
        return float(ms.value) if timed else None

    def _langevin(self, fn, nsteps, dt, dx, eta, mu, inv_a2, noise_scale, seed, step0, timed):
        p = StepParams(dt, dx, eta, mu, inv_a2, noise_scale)
        s0 = self.step_count if step0 is None else int(step0)
        ms = C.c_float(0.0)
        _lib.call(fn, self._h, C.byref(p), int(nsteps), s0 & 0xFFFFFFFFFFFFFFFF,
                  int(seed) & 0xFFFFFFFFFFFFFFFF, C.byref(ms) if timed else None)
        self.step_count = s0 + int(nsteps)
        return float(ms.value) if timed else None

    def step_overdamped(self, nsteps: int, *, dt: float, dx: float, eta_eff: float, mu: float,
                        inv_a2: float = 1.0, noise_scale: float = 0.0, seed: int = 0,
                        step0: Optional[int] = None, timed: bool = False) -> Optional[float]:
        """Advance nsteps of overdamped Euler-Maruyama (overdamped_euler_step_table);
        noise_scale = sqrt(2 T dt / (eta_eff mu^2 dx^3)); pi is left untouched."""
        return self._langevin("tip_lattice_step_overdamped", nsteps, dt, dx, eta_eff, mu,
                              inv_a2, noise_scale, seed, step0, timed)

    def step_baoab(self, nsteps: int, *, dt: float, dx: float, gamma: float, mu: float,
                   inv_a2: float = 1.0, noise_scale: float = 0.0, seed: int = 0,
                   step0: Optional[int] = None, timed: bool = False) -> Optional[float]:
        """Advance nsteps of BAOAB (baoab_step_table), one lattice pass per step;
        noise_scale = sqrt(T / (mu^2 dx^3)).  The last step's closing half-kick is deferred to
        the next BAOAB call or to the first operation that needs pi (download of a momentum,
        upload, set_table, set_zn, another stepper, finish())."""
        return self._langevin("tip_lattice_step_baoab", nsteps, dt, dx, gamma, mu, inv_a2,
                              noise_scale, seed, step0, timed)

    def finish(self) -> None:
        """Complete deferred work (the BAOAB closing half-kick) and synchronise."""
        _lib.call("tip_lattice_finish", self._h)

    def step_mlp(self, nsteps: int, surrogate, *, T_start: float, dT_per_step: float = 0.0,
                 dt: float, dx: float, eta_eff: float, mu: float, inv_a2: float = 1.0,
                 noise_scale: float = 0.0, seed: int = 0, step0: Optional[int] = None,
                 grid_points: int = 0, timed: bool = False) -> Optional[float]:
        """Complex field: advance nsteps with V' from a trained tip.surrogate.Surrogate at
        T = T_start + dT_per_step * n (continuous in T, no table of the potential).
        grid_points = 0 evaluates the network inline at every site; > 0 evaluates it once per
        step on that many points of its compressed rho coordinate and interpolates."""
        if not self.complex:
            raise TypeError("step_mlp needs a complex lattice")
        p = StepParams(dt, dx, eta_eff, mu, inv_a2, noise_scale)
        s0 = self.step_count if step0 is None else int(step0)
        ms = C.c_float(0.0)
        fm = surrogate.spec.features()
        _lib.call("tip_lattice_step_mlp", self._h, C.byref(p), surrogate.params.ptr, C.byref(fm),
                  float(T_start), float(dT_per_step), int(nsteps), s0 & 0xFFFFFFFFFFFFFFFF,
                  int(seed) & 0xFFFFFFFFFFFFFFFF, int(grid_points), C.byref(ms) if timed else None)
        self.step_count = s0 + int(nsteps)
        return float(ms.value) if timed else None

    # -- observables ---------------------------------------------------------------------
    def vacuum_count(self, esc: float = 1e4) -> int:
        c = C.c_ulonglong(0)
        _lib.call("tip_lattice_vacuum_count", self._h, float(esc), C.byref(c))
        return int(c.value)

    def vacuum_fraction(self, esc: float = 1e4) -> float:
        return self.vacuum_count(esc) / float(np.prod(self.shape))

    def winding_counts(self, plaq_ptr: Optional[int] = None, wsum_ptr: Optional[int] = None):
        """(string sites, nonzero plaquettes) of the current complex field; optionally the
        int8 [3][N] plaquette windings at plaq_ptr and the float32 per-site winding density
        (the reference's winding buffer) at wsum_ptr, both device pointers."""
        c = (C.c_ulonglong * 2)()
        _lib.call("tip_lattice_winding", self._h, plaq_ptr, wsum_ptr, c)
        return int(c[0]), int(c[1])

    def rho_theta(self, rho_ptr: Optional[int] = None, theta_ptr: Optional[int] = None):
        """Fill rho = |phi| and theta = arg(phi) (float32, formed in float64) at the device
        pointers given; returns (rho_min, rho_max) of the float64 rho."""
        mm = (C.c_double * 2)()
        _lib.call("tip_lattice_rho_theta", self._h, rho_ptr, theta_ptr, mm)
        return float(mm[0]), float(mm[1])

    def save_point(self) -> dict:
        """The complex driver's save-point quantities computed on the device
        (latticeSimComplex_numba.py:4625-4646): rho, theta, the winding buffer, rho_min,
        rho_max and the string count; only the finished float32 arrays cross PCIe."""
        if not self.complex:
            raise ValueError("save_point needs a complex lattice")
        from .device import DeviceArray
        rho = DeviceArray(self.shape, np.float32)
        theta = DeviceArray(self.shape, np.float32)
        wind = DeviceArray(self.shape, np.float32)
        try:
            lo, hi = self.rho_theta(rho.ptr, theta.ptr)
            nstr, _ = self.winding_counts(None, wind.ptr)
            return dict(rho=rho.download(), theta=theta.download(), winding=wind.download(),
                        rho_min=lo, rho_max=hi, n_strings=nstr)
        finally:
            for a in (rho, theta, wind):
                a.free()

    def minmax(self):
        out = (C.c_double * 2)()
        _lib.call("tip_lattice_minmax", self._h, out)
        return float(out[0]), float(out[1])

    def save_cosmolattice(self, path: str, *, step: int, t: float, T: float, a: float = 1.0,
                          H: float = 0.0, fStar: float = 1.0) -> None:
        """Write the current state as a CosmoLattice FLPI / FLP2 raw snapshot (program units
        = GeV / fStar; cubic lattices only)."""
        from .cosmolattice import write_raw_snapshot
        names = ["phi1", "phi2", "pi1", "pi2"] if self.complex else ["phi", "pi"]
        write_raw_snapshot(path, {k: self.download(k) for k in names}, step=step, t=t, T=T, a=a,
                           H=H, fStar=fStar)

    def load_cosmolattice(self, path: str) -> dict:
        """Upload a CosmoLattice raw snapshot (converted to GeV) into this lattice; returns its
        header values."""
        from .cosmolattice import read_raw_snapshot
        snap = read_raw_snapshot(path, derived=False)
        names = ["phi1", "phi2", "pi1", "pi2"] if self.complex else ["phi", "pi"]
        if snap["n_scalars"] != (2 if self.complex else 1):
            raise ValueError("snapshot and lattice differ in the number of scalars")
        for k in names:
            if k not in snap:
                raise ValueError(f"snapshot has no {k}")
            self.upload(k, snap[k].astype(self.dtype))
        return {k: v for k, v in snap.items() if not isinstance(v, np.ndarray)}

    def stream(self) -> int:
        """The lattice's HIP stream (hipStream_t as an integer)."""
        p = C.c_void_p()
        _lib.call("tip_lattice_stream", self._h, C.byref(p))
        return int(p.value or 0)

    def bubbles(self, esc: float = 1e4, dx_phys: Optional[float] = None,
                periodic: bool = False):
        """Connected regions of |phi| > esc of the current real field with the reference's
        per-bubble measures (tip.analysis.identify_bubbles), labelled on the device."""
        if self.complex:
            raise TypeError("bubbles: needs a real lattice")
        from .analysis import identify_bubbles
        from .device import DeviceArray
        v = DeviceArray(self.shape, self.dtype, ptr=self.field_ptr("phi"))
        _lib.call("tip_stream_synchronize", self.stream())  # steps in flight on the lattice
        return identify_bubbles(v, esc, dx_phys, periodic)

    def strings(self, threshold: float = 0.5, periodic: bool = False):
        """(labels, strings) of the current complex field: the winding buffer and its
        components (tip.analysis.identify_strings), both computed on the device."""
        if not self.complex:
            raise TypeError("strings: needs a complex lattice")
        from .analysis import identify_strings
        from .device import DeviceArray
        w = DeviceArray(self.shape, np.float32)
        try:
            self.winding_counts(None, w.ptr)
            return identify_strings(w, threshold, periodic)
        finally:
            w.free()

    def close(self) -> None:
        if getattr(self, "_h", None) and self._h.value:
            _lib.load().tip_lattice_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
