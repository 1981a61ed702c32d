"""Driver host logic (CPU): configuration defaults follow the reference drivers, the schedule
and the table-range rule."""
import math

import numpy as np

from tip import driver as D
from tip import potential as P


def test_defaults_follow_reference():
    r = D.RunConfig(field="real")
    assert r.dx_phys == 1e-3 and r.dt_phys == 1e-4 and r.eta_phys == r.T0  # :162-171, :646
    assert r.potential_type == "V_correct"
    c = D.RunConfig(field="complex")
    assert c.dt_phys == 2.5e-4 and c.eta_phys == 1000.0 and c.potential_type == "V_p"
    m = D.model_for(c)
    assert m.lam == 1e-16 and (m.coeff_b, m.coeff_f) == (2.0, -1.0)


def test_initial_fields_shapes_and_scale():
    f = D.initial_fields(D.RunConfig(field="real", Nx=8, Ny=8, Nz=8))
    assert f["phi"].shape == (8, 8, 8) and np.all(f["pi"] == 0)
    assert 0.005 < f["phi"].std() < 0.02
    c = D.initial_fields(D.RunConfig(field="complex", Nx=8, Ny=8, Nz=8))
    rho = np.hypot(c["phi1"], c["phi2"])
    assert rho.max() < 0.1 and np.all(c["pi1"] == 0)


def test_hubble_recurrence_closed_form():
    """The scale factor obeys a += a H dt/mu with H from T = T0/a (:3622-3637)."""
    h = P.HubbleState(T0=7000.0, mu=1000.0, eta=7.0)
    a = 1.0
    for _ in range(5):
        s = h.step(0.25)
        T = 7000.0 / a
        H = math.sqrt((T ** 4 * 106.75 * math.pi ** 2 / 30.0 + 1e36 / 4) / (3 * 2.4e18 ** 2))
        assert abs(s["T"] - T) < 1e-9 * T and abs(s["eta_eff"] - (7.0 + 3 * H / 1000.0)) < 1e-15
        a = a + a * H * 0.25 / 1000.0
        assert abs(h.a - a) < 1e-15


def test_noise_amplitude_formula():
    assert abs(P.noise_scale(7.0, 7000.0, 0.25, 1000.0, 0.002) - 1750.0) < 1e-9
