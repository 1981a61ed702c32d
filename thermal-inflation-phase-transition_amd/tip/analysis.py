"""Bubble and string analysis on the device: the connected-component statistics of
postprocess/revisualize_snapshots.py, with libtip's labelling in place of
scipy.ndimage.label.

  label(v, threshold, periodic=False) -> (labels, n)            scipy.ndimage.label(|v| > thr)
  identify_bubbles(phi, esc, dx_phys=None) -> [dict]             :171-213, mask :248-252
  identify_strings(winding, threshold=0.5) -> (labels, [dict])   :360-392
  bubble_summary(bubbles, lattice_volume) -> dict                :846-866

Fields may be numpy arrays (uploaded once) or tip.device.DeviceArray / (pointer, shape, dtype)
of a device-resident lattice, so a snapshot never has to leave the GPU to be analysed.  Labels
follow scipy's numbering (6-connectivity, no wrap-around, components 1..n in raster order of
their first site).  Per-component sums are exact integers for the coordinates; the value sum
(only its sign is used) is float64.
"""
from __future__ import annotations

import ctypes as C
import math
from typing import List, Optional, Tuple

import numpy as np

from . import _lib
from .device import DeviceArray


def _as_device(v) -> Tuple[DeviceArray, bool]:
    """(device array, owned) for a numpy array or a DeviceArray."""
    if isinstance(v, DeviceArray):
        return v, False
    a = np.ascontiguousarray(v)
    if a.dtype not in (np.float32, np.float64):
        a = a.astype(np.float64)
    if a.ndim != 3:
        raise ValueError(f"expected a 3-D field, got shape {a.shape}")
    return DeviceArray.from_numpy(a), True


def _suffix(dtype) -> str:
    dt = np.dtype(dtype)
    if dt == np.float32:
        return "f32"
    if dt == np.float64:
        return "f64"
    raise TypeError(f"unsupported dtype {dt}")


class Components:
    """Device labels of {|v| > threshold} and their per-component sums."""

    def __init__(self, v, threshold: float, periodic: bool = False, stats: bool = True):
        self.field, owned = _as_device(v)
        try:
            self.shape = self.field.shape
            sfx = _suffix(self.field.dtype)
            self.labels = DeviceArray(self.shape, np.int32)
            n = C.c_int64(0)
            _lib.call(f"tip_label_{sfx}", self.field.ptr, float(threshold), *self.shape,
                      int(bool(periodic)), self.labels.ptr, C.byref(n), None)
            self.n = int(n.value)
            self.volume = self.sum_xyz = self.sum_sq = self.sum_v = self.peak = None
            if stats and self.n > 0:
                nc = self.n
                d = {k: DeviceArray((m * nc,), dt) for k, m, dt in (
                    ("volume", 1, np.uint64), ("sum_xyz", 3, np.uint64),
                    ("sum_sq", 1, np.uint64), ("sum_v", 1, np.float64),
                    ("peak", 1, np.uint64))}
                _lib.call(f"tip_label_stats_{sfx}", self.field.ptr, self.labels.ptr, nc,
                          *self.shape, d["volume"].ptr, d["sum_xyz"].ptr, d["sum_sq"].ptr,
                          d["sum_v"].ptr, d["peak"].ptr, None)
                self.volume = d["volume"].download().astype(np.int64)
                self.sum_xyz = d["sum_xyz"].download().reshape(nc, 3)
                self.sum_sq = d["sum_sq"].download()
                self.sum_v = d["sum_v"].download()
                self.peak = d["peak"].download().view(np.float64)
                for a in d.values():
                    a.free()
        finally:
            if owned:
                self.field.free()
                self.field = None

    def labels_host(self) -> np.ndarray:
        return self.labels.download()

    def centroids(self) -> np.ndarray:
        """(n, 3) mean site coordinates: exact integer sums divided once (numpy's mean of the
        integer coordinates gives the same correctly rounded value while the sums stay below
        2^53)."""
        if self.n == 0:
            return np.zeros((0, 3))
        if int(self.sum_xyz.max()) < (1 << 53):
            return self.sum_xyz.astype(np.float64) / self.volume[:, None].astype(np.float64)
        return np.array([[int(self.sum_xyz[c, a]) / int(self.volume[c]) for a in range(3)]
                         for c in range(self.n)], dtype=np.float64)

    def r_rms(self) -> np.ndarray:
        """sqrt(mean |r - centroid|^2) per component from the exact integer numerator
        V sum|r|^2 - |sum r|^2 (int64 where it cannot overflow, Python integers otherwise)."""
        if self.n == 0:
            return np.zeros(0)
        v = self.volume.astype(np.int64)
        if int(v.max()) * int(self.sum_sq.max()) < (1 << 62):
            sx = self.sum_xyz.astype(np.int64)
            num = v * self.sum_sq.astype(np.int64) - (sx * sx).sum(axis=1)
            return np.sqrt(num.astype(np.float64) / (v.astype(np.float64) ** 2))
        out = np.empty(self.n)
        for c in range(self.n):
            vc = int(v[c])
            num = vc * int(self.sum_sq[c]) - sum(int(self.sum_xyz[c, a]) ** 2 for a in range(3))
            out[c] = math.sqrt(num / (vc * vc))
        return out

    def close(self) -> None:
        if getattr(self, "labels", None) is not None:
            self.labels.free()
            self.labels = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def label(v, threshold: float, periodic: bool = False):
    """(labels int32 array, number of components) of {|v| > threshold}, numbered as
    scipy.ndimage.label numbers them."""
    comp = Components(v, threshold, periodic, stats=False)
    try:
        return comp.labels_host(), comp.n
    finally:
        comp.close()


def identify_bubbles(phi, esc: float, dx_phys: Optional[float] = None,
                     periodic: bool = False) -> List[dict]:
    """Connected regions of |phi| > esc with the reference's per-bubble measures
    (revisualize_snapshots.py:171-213; the mask (phi > esc) | (phi < -esc) of :248-250),
    sorted by volume, largest first (ties in label order)."""
    comp = Components(phi, esc, periodic)
    try:
        out = []
        if comp.n == 0:
            return out
        cen, rr = comp.centroids(), comp.r_rms()
        vols = comp.volume.tolist()
        signs = np.where(comp.sum_v >= 0, 1, -1).tolist()
        peaks = comp.peak.tolist()
        rrl = rr.tolist()
        for c in range(comp.n):
            vol = vols[c]
            e = dict(centroid=cen[c], volume=vol,
                     r_eff=(3.0 * vol / (4.0 * np.pi)) ** (1.0 / 3.0), r_rms=rrl[c],
                     sign=signs[c], peak=peaks[c])
            if dx_phys is not None:
                e["r_phys"] = e["r_eff"] * dx_phys
            out.append(e)
        out.sort(key=lambda b: b["volume"], reverse=True)
        return out
    finally:
        comp.close()


def identify_strings(winding, threshold: float = 0.5, periodic: bool = False):
    """(labels, strings): connected string voxels |W| > threshold with loop_id, n_voxels,
    centroid, winding_sign and max_winding (revisualize_snapshots.py:360-392), sorted by
    n_voxels, longest first."""
    comp = Components(winding, threshold, periodic)
    try:
        strings = []
        cen = comp.centroids()
        vols = comp.volume.tolist() if comp.n else []
        signs = np.where(comp.sum_v >= 0, 1, -1).tolist() if comp.n else []
        peaks = comp.peak.tolist() if comp.n else []
        for c in range(comp.n):
            strings.append(dict(loop_id=c + 1, n_voxels=vols[c], centroid=cen[c],
                                winding_sign=signs[c], max_winding=peaks[c]))
        strings.sort(key=lambda s: s["n_voxels"], reverse=True)
        return comp.labels_host(), strings
    finally:
        comp.close()


def bubble_summary(bubbles: List[dict], lattice_volume: int) -> dict:
    """The per-snapshot summary row of revisualize_snapshots.py:846-866: counts by sign,
    largest and mean effective radius, total volume, volume fraction, and the percolation flag
    (volume fraction > 0.99999)."""
    n = len(bubbles)
    if n == 0:
        return dict(n_bub=0, n_pos=0, n_neg=0, r_max=0.0, r_mean=0.0, tot_vol=0, vol_frac=0.0,
                    percolated=False)
    n_pos = sum(1 for b in bubbles if b["sign"] > 0)
    r_all = [b["r_eff"] for b in bubbles]
    tot = sum(b["volume"] for b in bubbles)
    frac = tot / lattice_volume
    return dict(n_bub=n, n_pos=n_pos, n_neg=n - n_pos, r_max=max(r_all),
                r_mean=float(np.mean(r_all)), tot_vol=tot, vol_frac=frac,
                percolated=frac > 0.99999)
