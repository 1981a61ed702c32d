"""CosmoLattice interchange on the GPU (-m gpu): the port's hash initial conditions against the
oracle (cosmolattice_ext/models/thermal_inflation.h:174-222), a raw snapshot of a resident
lattice read back (FLPI / FLP2, field_snapshot.hpp:38-56) and the exporter's derived fields
(tools/export_cl_snapshots.py:87-106) computed on the device.

Bars: float64 initial field within 1e-15 relative of the oracle (libm log/cos), float32 within
one float32 ulp of the rounded oracle value; snapshot round trip exact in program units;
winding buffer equal to the oracle's rounded site sums; rho / theta equal numpy's."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n_scalars", [1, 2])
def test_hash_ics_match_oracle(n_scalars):
    from tip import cosmolattice as CL
    shape, fstar = (24, 20, 36), 3.7
    amp = (0.01 / fstar) / np.sqrt(2.0) if n_scalars == 2 else 0.01 / fstar
    r0, r1 = O.cl_hash_ics(shape, amp, two=n_scalars == 2)
    g64 = CL.hash_ics(shape, n_scalars, fstar, np.float64)
    g32 = CL.hash_ics(shape, n_scalars, fstar, np.float32)
    for g, r in zip(g64, (r0, r1)):
        assert np.max(np.abs(g - r) / np.abs(r)) < 1e-15
    for g, r in zip(g32, (r0, r1)):
        ulp = np.abs(g.view(np.int32).astype(np.int64) - r.astype(np.float32).view(np.int32))
        assert int(ulp.max()) <= 1


def test_lattice_snapshot_round_trip(tmp_path):
    from tip import cosmolattice as CL
    from tip.lattice import Lattice
    n, fstar = 32, 2.5
    rng = np.random.default_rng(4)
    init = {k: (rng.normal(size=(n, n, n)) * 300.0).astype(np.float32)
            for k in ("phi1", "phi2", "pi1", "pi2")}
    lat = Lattice((n, n, n), np.float32, complex_field=True)
    try:
        for k, v in init.items():
            lat.upload(k, v)
        path = str(tmp_path / CL.snapshot_name(40))
        lat.save_cosmolattice(path, step=40, t=0.01, T=7100.0, a=1.0, H=0.0, fStar=fstar)
        snap = CL.read_raw_snapshot(path)
        prog = CL.read_raw_snapshot(path, units="program", derived=False)
        for k, v in init.items():
            assert np.array_equal(prog[k], (v.astype(np.float64) / fstar).astype(np.float32))
        p1, p2 = snap["phi1"], snap["phi2"]
        w_ref, _ = O.winding(p1, p2)
        assert np.array_equal(snap["winding"], np.rint(w_ref).astype(np.float32))
        assert np.array_equal(snap["rho"], np.sqrt(p1 ** 2 + p2 ** 2))
        out = tmp_path / "state.npz"
        CL.export_npz(snap, str(out))
        with np.load(out) as z:
            assert {"phi1", "phi2", "rho", "theta", "winding", "pi1", "pi2", "rho_min",
                    "rho_max", "scale_factor", "hubble"} <= set(z.files)
        lat2 = Lattice((n, n, n), np.float32, complex_field=True)
        try:
            hdr = lat2.load_cosmolattice(path)
            assert hdr["step"] == 40 and hdr["temperature"] == 7100.0
            assert np.array_equal(lat2.download("phi1"), p1.astype(np.float32))
        finally:
            lat2.close()
    finally:
        lat.close()
