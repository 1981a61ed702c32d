"""The C-ABI library loads and exports exactly what include/tip.h declares (CPU, no compute)."""
import ctypes as C
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


def _declared():
    text = (ROOT / "include" / "tip.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(tip_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    import tip
    lib = tip.load()
    names = _declared()
    assert len(names) > 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_ctypes_table_covers_header():
    from tip import _lib
    assert sorted(_lib.SIGNATURES) == _declared()


def test_argument_errors_without_compute():
    from tip import _lib
    lib = _lib.load()
    # bad dims are rejected on the host before any HIP call
    rc = lib.tip_rk2_fused_f64(None, None, None, None, 0.25, 2.0, 7.0, 1000.0, 1.0, None, None,
                               0.0, 1.0, 16, 0, 0, 4, 4, 0, 0.0, None, 0.0, None)
    assert rc == -1
    assert "lattice dims" in _lib.last_error()
    rc = lib.tip_rk2_fused_f32(None, None, None, None, 0.25, 2.0, 7.0, 1000.0, 1.0, None, None,
                               0.0, 1.0, 16, 0, 70000, 4, 4, 0, 0.0, None, 0.0, None)
    assert rc == -1
    assert "exceeds" in _lib.last_error()
    rc = lib.tip_winding_f32(None, None, None, None, None, 4, 4, 4, None)
    assert rc == -1
    assert lib.tip_noise_seed(5, 0) == 5


def test_drop_in_rejects_bad_dtypes():
    import numpy as np
    from tip import kernels
    phi = np.zeros((4, 4, 4))
    with pytest.raises(TypeError):
        kernels.rk2_step_table(phi, np.zeros((4, 4, 4), np.float32), 0.25, 2.0, 7.0, 1000.0, None,
                               phi.copy(), phi.copy(), 1.0, np.zeros(8), 0.0, 1.0, 8)
    with pytest.raises(ValueError):
        kernels.rk2_step_table(phi, np.zeros((4, 4, 5)), 0.25, 2.0, 7.0, 1000.0, None,
                               phi.copy(), phi.copy(), 1.0, np.zeros(8), 0.0, 1.0, 8)
    with pytest.raises(TypeError):
        kernels.rk2_step_table(np.zeros((4, 4)), np.zeros((4, 4)), 0.25, 2.0, 7.0, 1000.0, None,
                               None, None, 1.0, np.zeros(8), 0.0, 1.0, 8)
