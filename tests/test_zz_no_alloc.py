"""The allocation-free float32 drop-in (tip_rk2_step_table_f32_noalloc) against the allocating
one (tip_rk2_step_table_f32): identical fields and vacuum counts, called through the C-ABI on
device buffers, twice in a row with the second call on a smaller lattice and table."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _table(npts, slope, seed):
    x = np.linspace(0.0, 3.0e4, npts)
    rng = np.random.default_rng(seed)
    return slope * x + 1.0e3 * rng.standard_normal(npts), 0.0, (npts - 1) / 3.0e4


def _one_call(name, shape, phi0, pi0, table, tmin, dinv, step):
    from tip import _lib
    from tip.device import DeviceArray

    phi, pi = DeviceArray.from_numpy(phi0), DeviceArray.from_numpy(pi0)
    pt, qt = DeviceArray(shape, np.float32), DeviceArray(shape, np.float32)
    tab = DeviceArray.from_numpy(table)
    vac = DeviceArray.zeros((1,), np.uint64)
    nx, ny, nz = shape
    _lib.call(name, phi.ptr, pi.ptr, 0.25, 2.0, 7.0, 1000.0, None, pt.ptr, qt.ptr, 1.0, tab.ptr,
              tmin, dinv, len(table), 1, nx, ny, nz, step, 1750.0, vac.ptr, 1.0e2, None)
    out = phi.download(), pi.download(), int(vac.download()[0])
    for a in (phi, pi, pt, qt, tab, vac):
        a.free()
    return out


@pytest.mark.parametrize("calls", [[((32, 24, 64), 8192, 11), ((16, 8, 40), 513, 12)]])
def test_noalloc_matches_allocating_drop_in(calls):
    rng = np.random.default_rng(7)
    for shape, npts, step in calls:
        phi0 = (300.0 * rng.standard_normal(shape)).astype(np.float32)
        pi0 = (50.0 * rng.standard_normal(shape)).astype(np.float32)
        table, tmin, dinv = _table(npts, -1.0e6, step)
        a = _one_call("tip_rk2_step_table_f32", shape, phi0, pi0, table, tmin, dinv, step)
        b = _one_call("tip_rk2_step_table_f32_noalloc", shape, phi0, pi0, table, tmin, dinv,
                      step)
        print(f"no-alloc step {shape} npts={npts}: vac {a[2]} / {b[2]}", flush=True)
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
        assert a[2] == b[2]
        assert not np.array_equal(a[0], phi0)


def test_noalloc_rejects_oversized_table():
    from tip import _lib

    lib = _lib.load()
    rc = lib.tip_rk2_step_table_f32_noalloc(
        C.c_void_p(16), C.c_void_p(16), 0.25, 2.0, 7.0, 1000.0, None, C.c_void_p(16),
        C.c_void_p(16), 1.0, C.c_void_p(16), 0.0, 1.0, 65538, 1, 4, 4, 4, 0, 0.0, None, 0.0, None)
    assert rc == -1
    assert "static table scratch" in _lib.last_error()
