"""Batched ensemble of independent real-field realisations (one launch per step, hipGraph
replay of K-step loops).  The reference runs one realisation per process; nucleation-rate
statistics need many, each following rk2_step_table
(simulation/latticeSimeRescale_numba.py:1875-1966) with its own seed, T0 and V' table.
"""
from __future__ import annotations

import ctypes as C
from typing import Optional, Sequence

import numpy as np

from . import _lib
from ._lib import StepParams

SEQUENTIAL, BATCHED, GRAPH = 0, 1, 2
MODES = {"sequential": SEQUENTIAL, "batched": BATCHED, "graph": GRAPH}


class Ensemble:
    def __init__(self, shapes: Sequence[Sequence[int]], seeds: Optional[Sequence[int]] = None):
        self.shapes = [tuple(int(v) for v in s) for s in shapes]
        n = len(self.shapes)
        dims = np.ascontiguousarray(np.array(self.shapes, dtype=np.int64).reshape(-1))
        sd = np.ascontiguousarray(
            np.arange(n, dtype=np.uint64) if seeds is None else np.array(seeds, dtype=np.uint64))
        h = C.c_void_p()
        _lib.call("tip_ensemble_create", C.byref(h), n, dims.ctypes.data, sd.ctypes.data)
        self._h = h
        self.n = n
        self.sites = np.array([int(np.prod(s)) for s in self.shapes], dtype=np.int64)

    def upload(self, i: int, name: str, x: np.ndarray) -> None:
        x = np.ascontiguousarray(x, dtype=np.float32)
        if x.shape != self.shapes[i]:
            raise ValueError(f"instance {i}: shape {x.shape} != {self.shapes[i]}")
        _lib.call("tip_ensemble_upload", self._h, i, ["phi", "pi"].index(name), x.ctypes.data)

    def download(self, i: int, name: str) -> np.ndarray:
        out = np.empty(self.shapes[i], np.float32)
        _lib.call("tip_ensemble_download", self._h, i, ["phi", "pi"].index(name), out.ctypes.data)
        return out

    def set_table(self, i: int, table, tmin: float, dinv: float, odd: bool = False) -> None:
        t = np.ascontiguousarray(table, dtype=np.float64)
        _lib.call("tip_ensemble_set_table", self._h, i, t.ctypes.data, t.size, float(tmin),
                  float(dinv), int(bool(odd)))

    def set_params(self, i: int, *, dt, dx, eta_eff, mu, inv_a2=1.0, noise_scale=0.0) -> None:
        p = StepParams(dt, dx, eta_eff, mu, inv_a2, noise_scale)
        _lib.call("tip_ensemble_set_params", self._h, i, C.byref(p))

    def set_geometry(self, rows=0, planes_per_block=8, xcd_remap=True, persistent_blocks=0,
                     kernel=0) -> None:
        """kernel 0 = auto (LDS row tiles where every nz <= 256), 1 = row tiles (`rows` = y rows
        per thread), 2 = flat tiles (`rows` = rows of 64 sites per block)."""
        _lib.call("tip_ensemble_set_geometry", self._h, int(kernel), int(rows),
                  int(planes_per_block), int(bool(xcd_remap)), int(persistent_blocks))

    def set_schedule(self, group_bytes: int = 0) -> None:
        """Step instances in groups of about group_bytes of field data (16 B per site; each
        block-size class of the row kernel cut into groups of equal bytes), each group through
        all steps of a call before the next (0 = one group per class)."""
        _lib.call("tip_ensemble_set_schedule", self._h, int(group_bytes))

    def set_concurrency(self, groups_at_once: int = 1) -> None:
        """1 or 2 groups of the schedule stepping at once (two streams); same results."""
        _lib.call("tip_ensemble_set_concurrency", self._h, int(groups_at_once))

    def step(self, nsteps: int, mode: str = "graph", timed: bool = False):
        ms = C.c_float(0.0)
        _lib.call("tip_ensemble_step", self._h, int(nsteps), MODES[mode],
                  C.byref(ms) if timed else None)
        return float(ms.value) if timed else None

    def vacuum_counts(self, esc: float = 1e4) -> np.ndarray:
        out = np.zeros(self.n, np.uint64)
        _lib.call("tip_ensemble_vacuum_counts", self._h, float(esc), out.ctypes.data)
        return out

    def vacuum_fractions(self, esc: float = 1e4) -> np.ndarray:
        return self.vacuum_counts(esc).astype(np.float64) / self.sites

    @property
    def step_count(self) -> int:
        v = C.c_uint64(0)
        _lib.call("tip_ensemble_step_count", self._h, C.byref(v))
        return int(v.value)

    def set_step(self, step: int) -> None:
        _lib.call("tip_ensemble_set_step", self._h, int(step))

    def close(self) -> None:
        if getattr(self, "_h", None) and self._h.value:
            _lib.load().tip_ensemble_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def nucleation_ensemble_shapes():
    """64 realisations of uneven size: 32 x 64^3, 24 x 96^3, 8 x 128^3 (46,399,488 sites)."""
    return [(64,) * 3] * 32 + [(96,) * 3] * 24 + [(128,) * 3] * 8
