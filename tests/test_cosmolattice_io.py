"""CosmoLattice raw snapshot I/O and the port's hash initial conditions (CPU).

Pins: the header is the packed layout of cosmolattice_ext/measurements/field_snapshot.hpp:
43-56 (56 / 60 bytes, the exporter's struct strings "<IIq5d" / "<IIq5dI",
tools/export_cl_snapshots.py:36-39), magics 'FLPI' / 'FLP2' as little-endian bytes, payload
order phi [phi2] pi [pi2]; a file assembled byte by byte from those formats reads back; the
oracle's hash initial conditions equal a pure-Python restatement of thermal_inflation.h:
188-227 and the splitmix64 reference output, with unit Gaussian moments."""
import math
import struct

import numpy as np
import pytest

from oracle import oracle as O


def _cl():
    from tip import cosmolattice as CL
    return CL


def test_header_layout_and_magic(tmp_path):
    CL = _cl()
    assert struct.calcsize(CL.HEADER_FMT) == 56 and struct.calcsize(CL.HEADER2_FMT) == 60
    n = 4
    rng = np.random.default_rng(0)
    f = {k: rng.normal(size=(n, n, n)).astype(np.float32) for k in ("phi", "pi")}
    p = tmp_path / "a.raw"
    CL.write_raw_snapshot(str(p), f, step=12, t=3.5, T=7000.0, a=1.25, H=2e-3, fStar=2.0,
                          units="program")
    b = p.read_bytes()
    assert b[:4] == b"IPLF" and len(b) == 56 + 2 * 4 * n ** 3
    assert struct.unpack("<IIq5d", b[:56]) == (0x464C5049, n, 12, 3.5, 7000.0, 1.25, 2e-3, 2.0)
    assert np.array_equal(np.frombuffer(b[56:56 + 4 * n ** 3], "<f4").reshape(n, n, n), f["phi"])
    s = CL.read_raw_snapshot(str(p))
    assert s["step"] == 12 and s["N"] == n and s["a"] == 1.25
    assert np.array_equal(s["phi"], f["phi"].astype(np.float64) * 2.0)
    assert np.array_equal(s["pi"], f["pi"].astype(np.float64) * 2.0)


def test_two_component_file_built_from_the_formats(tmp_path):
    """A file assembled from the exporter's struct format reads back field by field."""
    CL = _cl()
    n = 3
    rng = np.random.default_rng(1)
    arrs = [rng.normal(size=(n, n, n)).astype(np.float32) for _ in range(4)]
    raw = struct.pack("<IIq5dI", 0x464C5032, n, 7, 1.0, 6900.0, 1.0, 0.0, 0.5, 2)
    raw += b"".join(a.astype("<f4").tobytes() for a in arrs)
    p = tmp_path / "b.raw"
    p.write_bytes(raw)
    s = CL.read_raw_snapshot(str(p), units="program", derived=False)
    assert raw[:4] == b"2PLF" and s["n_scalars"] == 2 and s["temperature"] == 6900.0
    for k, a in zip(("phi1", "phi2", "pi1", "pi2"), arrs):
        assert np.array_equal(s[k], a)
    # round trip through the writer in GeV
    q = tmp_path / "c.raw"
    CL.write_raw_snapshot(str(q), {k: s[k].astype(np.float64) * 0.5 for k in s
                                   if k in ("phi1", "phi2", "pi1", "pi2")},
                          step=7, t=1.0, T=6900.0, fStar=0.5)
    assert q.read_bytes() == raw


def test_truncated_and_bad_magic(tmp_path):
    CL = _cl()
    p = tmp_path / "bad.raw"
    p.write_bytes(struct.pack("<IIq5d", 0x12345678, 2, 0, 0, 0, 0, 0, 1.0) + b"\0" * 64)
    with pytest.raises(ValueError):
        CL.read_raw_snapshot(str(p))
    p.write_bytes(struct.pack("<IIq5d", 0x464C5049, 2, 0, 0, 0, 0, 0, 1.0) + b"\0" * 10)
    with pytest.raises(ValueError):
        CL.read_raw_snapshot(str(p))


def test_manifest_round_trip(tmp_path):
    CL = _cl()
    d = str(tmp_path)
    CL.append_manifest(d, step=0, t=0.0, T=7350.0, a=1.0, H=1e-3, fStar=2.0, n_scalars=2)
    CL.append_manifest(d, step=100, t=0.025, T=7340.5, a=1.01, H=1e-3, fStar=2.0, n_scalars=2)
    CL.append_manifest(d, step=100, t=0.025, T=7340.5, a=1.01, H=1e-3, fStar=2.0, n_scalars=2)
    text = (tmp_path / "manifest.csv").read_text().splitlines()
    assert text[0] == "step,t,T,a,H,fStar,n_scalars,expansion_stage,rho_m,filename"
    rows = CL.read_manifest(d)
    assert [r["step"] for r in rows] == [0, 100]
    assert rows[1]["filename"] == "snapshot_0000000100.raw" and rows[1]["T"] == 7340.5


# ----------------------------------------------------------------------------------------
# hash initial conditions
# ----------------------------------------------------------------------------------------
M64 = (1 << 64) - 1


def mix64(x):
    x ^= x >> 30
    x = (x * 0xBF58476D1CE4E5B9) & M64
    x ^= x >> 27
    x = (x * 0x94D049BB133111EB) & M64
    return x ^ (x >> 31)


def hash_gaussian(seed):
    h1, h2 = mix64(seed), mix64(seed ^ 0x9E3779B97F4A7C15)
    u1 = ((h1 >> 11) + 1.0) * (1.0 / 9007199254740992.0)
    u2 = (h2 >> 11) * (1.0 / 9007199254740992.0)
    return math.sqrt(-2.0 * math.log(max(u1, 1e-300))) * math.cos(6.28318530717958647692 * u2)


def test_mixer_is_splitmix64():
    # splitmix64 from state 0: the first output is the finaliser of the golden increment
    assert mix64(0x9E3779B97F4A7C15) == 0xE220A8397B1DCDAF


def test_oracle_ics_equal_restatement():
    shape = (3, 4, 5)
    f0, f1 = O.cl_hash_ics(shape, 0.7, two=True)
    ic = 0x4E554D42412D4943
    for s in range(60):
        i, j, k = np.unravel_index(s, shape)
        assert f0[i, j, k] == pytest.approx(0.7 * hash_gaussian(s ^ ic), rel=1e-15)
        assert f1[i, j, k] == pytest.approx(0.7 * hash_gaussian(s ^ 0x9E3779B97F4A7C15 ^ ic),
                                            rel=1e-15)


def test_oracle_ics_moments():
    f0, f1 = O.cl_hash_ics((64, 64, 64), 1.0, two=True)
    for f in (f0, f1):
        assert abs(f.mean()) < 5e-3 and abs(f.std() - 1.0) < 5e-3
    assert abs(np.corrcoef(f0.ravel(), f1.ravel())[0, 1]) < 5e-3
