"""CosmoLattice raw snapshot I/O and initial conditions, for cross-checking runs here against
runs of the reference's CosmoLattice port.

  write_raw_snapshot / read_raw_snapshot   FLPI / FLP2 files of
      cosmolattice_ext/measurements/field_snapshot.hpp:38-56 (packed header, then float32
      N^3 payloads phi [phi2] pi [pi2] in program units, value = GeV / fStar); the reader
      returns the dict of tools/export_cl_snapshots.py:43-122
  append_manifest / read_manifest          field_states/manifest.csv (field_snapshot.hpp:296-321)
  export_npz                               the numba-driver npz the exporter writes
                                           (export_cl_snapshots.py:236-275); winding and
                                           rho/theta computed on the GPU
  hash_ics                                 the port's numba-compatible white-noise start
                                           (cosmolattice_ext/models/thermal_inflation.h:
                                           174-222), generated on the GPU
  Lattice.save_cosmolattice / load_cosmolattice (tip.lattice) wrap these for a resident
  lattice.

The files are little-endian.  Header formats (the exporter's struct strings):
  FLPI / FLPH  "<IIq5d"  magic, N, step, t, T, a, H, fStar            (56 bytes)
  FLP2         "<IIq5dI" ... , nScalars                               (60 bytes)
"""
from __future__ import annotations

import os
import struct
from typing import Dict, List, Optional

import numpy as np

from . import _lib
from .device import DeviceArray

MAGIC_PHI = 0x464C5048   # 'FLPH' phi only (legacy)
MAGIC_PI = 0x464C5049    # 'FLPI' phi + pi
MAGIC_PI2 = 0x464C5032   # 'FLP2' phi1, phi2, pi1, pi2
HEADER_FMT = "<IIq5d"
HEADER2_FMT = "<IIq5dI"
MANIFEST_HEADER = "step,t,T,a,H,fStar,n_scalars,expansion_stage,rho_m,filename\n"


def snapshot_name(step: int) -> str:
    return f"snapshot_{int(step):010d}.raw"


def write_raw_snapshot(path: str, fields: Dict[str, np.ndarray], *, step: int, t: float,
                       T: float, a: float = 1.0, H: float = 0.0, fStar: float = 1.0,
                       units: str = "GeV") -> None:
    """Write an FLPI (fields phi, pi) or FLP2 (phi1, phi2, pi1, pi2) snapshot.  units="GeV"
    divides by fStar in float64 before the float32 store (the port stores program units);
    units="program" writes the arrays as given."""
    two = "phi1" in fields
    names = ["phi1", "phi2", "pi1", "pi2"] if two else ["phi", "pi"]
    arrs = [np.asarray(fields[k]) for k in names]
    n = arrs[0].shape[0]
    for a_ in arrs:
        if a_.shape != (n, n, n):
            raise ValueError(f"CosmoLattice snapshots are cubic: got {a_.shape}")
    if units == "GeV":
        arrs = [(x.astype(np.float64) / fStar).astype(np.float32) for x in arrs]
    elif units == "program":
        arrs = [np.ascontiguousarray(x, dtype=np.float32) for x in arrs]
    else:
        raise ValueError(f"unknown units {units!r}")
    with open(path, "wb") as f:
        if two:
            f.write(struct.pack(HEADER2_FMT, MAGIC_PI2, n, int(step), float(t), float(T),
                                float(a), float(H), float(fStar), 2))
        else:
            f.write(struct.pack(HEADER_FMT, MAGIC_PI, n, int(step), float(t), float(T), float(a),
                                float(H), float(fStar)))
        for x in arrs:
            f.write(np.ascontiguousarray(x, dtype="<f4").tobytes())


# Payload order of each magic, as FieldSnapshotWriter lays the fields out after the packed
# header (field_snapshot.hpp:38-56): the header format and the field names in file order.
_LAYOUT = {
    MAGIC_PHI: (HEADER_FMT, ("phi",)),
    MAGIC_PI: (HEADER_FMT, ("phi", "pi")),
    MAGIC_PI2: (HEADER2_FMT, ("phi1", "phi2", "pi1", "pi2")),
}


def read_raw_snapshot(path: str, units: str = "GeV", derived: bool = True) -> dict:
    """Read an FLPH / FLPI / FLP2 snapshot into a dict with the exporter's keys: the fields in
    GeV as float64 (program-unit float32 with units="program") and the header scalars.  For
    two components and derived=True, rho, theta and the winding buffer are added; the
    winding is computed on the GPU from the float64 fields.

    The whole file is read once; its length must equal header + fields x 4 N^3 bytes exactly,
    so a short or over-long file is rejected before any field is decoded."""
    if units not in ("GeV", "program"):
        raise ValueError(f"unknown units {units!r}")
    blob = open(path, "rb").read()
    if len(blob) < 4:
        raise ValueError(f"{path}: {len(blob)} bytes, too short for a snapshot magic")
    (magic,) = struct.unpack_from("<I", blob, 0)
    if magic not in _LAYOUT:
        raise ValueError(f"{path}: magic {magic:#010x} is not FLPH, FLPI or FLP2")
    fmt, names = _LAYOUT[magic]
    hsize = struct.calcsize(fmt)
    if len(blob) < hsize:
        raise ValueError(f"{path}: {len(blob)} bytes, header needs {hsize}")
    head = struct.unpack_from(fmt, blob, 0)
    n, step, t, T, a, H, fStar = head[1:8]
    n3 = int(n) ** 3
    want = hsize + 4 * n3 * len(names)
    if len(blob) != want:
        raise ValueError(f"{path}: {len(blob)} bytes, N={n} with {len(names)} fields needs "
                         f"{want}")
    data = np.frombuffer(blob, dtype="<f4", count=n3 * len(names), offset=hsize)
    data = data.reshape((len(names), n, n, n))
    out = dict(step=int(step), time=float(t), temperature=float(T), a=float(a), H=float(H),
               fStar=float(fStar), N=int(n), n_scalars=2 if magic == MAGIC_PI2 else 1)
    for i, name in enumerate(names):
        out[name] = (data[i].astype(np.float64) * fStar if units == "GeV"
                     else data[i].astype(np.float32))
    if magic == MAGIC_PI2 and derived and units == "GeV":
        from .kernels_complex import compute_winding_number
        p1, p2 = out["phi1"], out["phi2"]
        out["rho"] = np.sqrt(p1 ** 2 + p2 ** 2)
        out["theta"] = np.arctan2(p2, p1)
        out["winding"] = compute_winding_number(p1, p2)[0]
    return out


def export_npz(snap: dict, path: str, hubble: bool = True) -> None:
    """The numba-driver npz of a read snapshot (keys of export_cl_snapshots.py:236-276;
    scale_factor and hubble unless hubble=False)."""
    if snap.get("n_scalars", 1) >= 2:
        rho = snap["rho"]
        d = dict(phi1=snap["phi1"].astype(np.float32), phi2=snap["phi2"].astype(np.float32),
                 rho=rho.astype(np.float32), theta=snap["theta"].astype(np.float32),
                 winding=np.asarray(snap["winding"], np.float32), step=snap["step"],
                 time=snap["time"], temperature=snap["temperature"],
                 rho_min=float(rho.min()), rho_max=float(rho.max()))
        if "pi1" in snap:
            d["pi1"] = snap["pi1"].astype(np.float32)
            d["pi2"] = snap["pi2"].astype(np.float32)
    else:
        phi = snap["phi"]
        d = dict(phi=phi.astype(np.float32), step=snap["step"], time=snap["time"],
                 temperature=snap["temperature"], phi_min=float(phi.min()),
                 phi_max=float(phi.max()))
        if "pi" in snap:
            d["pi"] = snap["pi"].astype(np.float32)
    if hubble:
        d["scale_factor"] = snap["a"]
        d["hubble"] = snap["H"]
    np.savez_compressed(path, **d)


def append_manifest(state_dir: str, *, step: int, t: float, T: float, a: float, H: float,
                    fStar: float, n_scalars: int, expansion_stage: int = 0,
                    rho_m: float = 0.0, filename: Optional[str] = None) -> None:
    """One manifest.csv row (the header is written with the file)."""
    os.makedirs(state_dir, exist_ok=True)
    path = os.path.join(state_dir, "manifest.csv")
    new = not os.path.exists(path)
    with open(path, "a") as f:
        if new:
            f.write(MANIFEST_HEADER)
        f.write(f"{int(step)},{t!r},{T!r},{a!r},{H!r},{fStar!r},{int(n_scalars)},"
                f"{int(expansion_stage)},{rho_m!r},{filename or snapshot_name(step)}\n")


def read_manifest(state_dir: str) -> List[dict]:
    """Rows of manifest.csv, first occurrence of each step (export_cl_snapshots.py:180-216)."""
    rows, seen = [], set()
    keys = ("step", "t", "T", "a", "H", "fStar", "n_scalars")
    with open(os.path.join(state_dir, "manifest.csv")) as f:
        for line in f:
            line = line.strip()
            if not line or line.startswith("step,"):
                continue
            parts = [p.strip() for p in line.split(",")]
            if len(parts) < 8:
                continue
            r = dict(zip(keys, parts))
            r["filename"] = parts[-1]
            step = int(float(r["step"]))
            if step in seen:
                continue
            seen.add(step)
            rows.append(dict(step=step, t=float(r["t"]), T=float(r["T"]), a=float(r["a"]),
                             H=float(r["H"]), fStar=float(r["fStar"]),
                             n_scalars=int(float(r["n_scalars"])), filename=r["filename"]))
    return rows


def hash_ics(shape, n_scalars: int = 1, fStar: float = 1.0, dtype=np.float64,
             units: str = "program"):
    """The port's numba-compatible initial field on the GPU: amplitude 0.01 GeV (0.01/sqrt 2
    per component for two scalars) times the site hash's Gaussian.  units="program" returns
    what the port holds (amp / fStar); "GeV" multiplies back by fStar in float64.  Returns
    a list of n_scalars arrays (pi starts at zero)."""
    nx, ny, nz = (int(s) for s in shape)
    amp = (0.01 / fStar) / np.sqrt(2.0) if n_scalars >= 2 else 0.01 / fStar
    dt = np.dtype(dtype)
    if units == "GeV":
        dt = np.dtype(np.float64)
    sfx = "f64" if dt == np.float64 else "f32"
    d0 = DeviceArray(shape, dt)
    d1 = DeviceArray(shape, dt) if n_scalars >= 2 else None
    try:
        _lib.call(f"tip_cl_hash_ics_{sfx}", d0.ptr, d1.ptr if d1 is not None else None, nx, ny,
                  nz, float(amp), None)
        out = [d0.download()] + ([d1.download()] if d1 is not None else [])
    finally:
        d0.free()
        if d1 is not None:
            d1.free()
    if units == "GeV":
        out = [(x * fStar).astype(np.dtype(dtype)) for x in out]
    return out
