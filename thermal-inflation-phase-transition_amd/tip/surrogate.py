"""Learned V'(rho, T) surrogate: data generation, GPU training, evaluation, on-device
re-tabulation and the inline-MLP complex step.

Ground truth is the float64 finite-temperature potential with the Coleman-Weinberg term
(tip.potential.vprime_full, the assembly of cosmolattice_ext/models/thermal_tables.hpp:96-146
with exact Bessel-series thermal functions).  With g(rho, T) = (V'/rho - (lam rho^2 - m^2)) /
(out_scale e(rho)) the thermal + CW part of V'/rho over the fixed envelope
e = 1/(1 + (rho/rho_cut)^2), the kernels assemble

    V'/rho = lam rho^2 - m^2 + out_scale e(rho) (g0(f2) + f1 y),

where g0(f2) is a cubic in the temperature feature fitted to g(0, T) (the effective mass
M_eff^2 = V'/rho at rho -> 0, which decides the spinodal instability near T_c and which the
network alone resolved only to ~5e-4 m^2) and y is the network (2 -> 32 -> 32 -> 1 tanh,
float32), which learns (g - g0) / f1.  f1 = 0 at rho = 0, so the effective mass is the fit's,
independent of the network; the tree term and the envelope are exact, so the network's
residual far above the thermal scale (where the true thermal part is exponentially small)
cannot leak into V' near the vacuum.

Sampling: the |phi| range [0, 1.5 phi_true] is ~1e11 GeV wide while the thermal barrier lives
at ~1e3 GeV, so half of the samples are uniform in |phi|/phi_true (the specified range) and
half uniform in the network's log-compressed input f1 = asinh(rho/rho_s)/asinh(rho_max/rho_s),
which is what lets 32 tanh units resolve the barrier.

Accuracy metric: |dV'| / max(|V'|, 1e-3 T^4/(2 pi^2) n) as in tools/validate_langevin.py:137-141
(n = the largest thermal multiplicity, at least 1).
"""
from __future__ import annotations

import ctypes as C
import dataclasses
import json
import math
import os
from dataclasses import dataclass, field
from typing import Dict, Optional

import numpy as np

from . import _lib
from . import potential as P
from ._lib import MlpFeatures
from .device import DeviceArray

H = 32
NPARAMS = 1185


@dataclass
class SurrogateSpec:
    model: P.ModelParams = field(default_factory=P.ModelParams.complex)
    T_lo: float = 6500.0
    T_hi: float = 7500.0
    T_ref: float = 7350.0      # temperature defining phi_true
    rho_frac: float = 1.5      # rho in [0, rho_frac * phi_true]
    rho_scale: float = 100.0   # GeV, width of the log-compression of rho
    rho_cut: float = 7.5e4     # GeV, envelope scale of the learned part (about 10 T)
    include_cw: bool = True

    def __post_init__(self):
        self.phi_true = P.thermal_vev(self.T_ref, self.model)
        self.rho_max = self.rho_frac * self.phi_true
        self.out_scale = self.model.mphi ** 2
        self.g0 = self.fit_g0()

    # rho at which V'/rho is taken as its rho -> 0 limit (its rho^2 term is < 1e-12 there)
    RHO_ZERO = 1e-2

    def fit_g0(self) -> np.ndarray:
        """Cubic in f2 through g(0, T) over the training range widened by 10 % on each side
        (least squares on 801 points; residual ~1e-8, against M_eff^2 / m^2 ~ 1e-3 at the
        top of the range)."""
        w = 0.1 * (self.T_hi - self.T_lo)
        T = np.linspace(self.T_lo - w, self.T_hi + w, 801)
        g = self.g_full(np.full_like(T, self.RHO_ZERO), T)
        f2 = (T - 0.5 * (self.T_lo + self.T_hi)) / (0.5 * (self.T_hi - self.T_lo))
        return np.polynomial.polynomial.polyfit(f2, g, 3)

    def g0_of(self, T) -> np.ndarray:
        f2 = (np.asarray(T, dtype=np.float64) - 0.5 * (self.T_lo + self.T_hi)) / (
            0.5 * (self.T_hi - self.T_lo))
        return np.polynomial.polynomial.polyval(f2, self.g0)

    def features(self) -> MlpFeatures:
        return MlpFeatures(self.rho_scale, self.rho_max, 0.5 * (self.T_lo + self.T_hi),
                           0.5 * (self.T_hi - self.T_lo), self.out_scale, self.model.lam,
                           self.model.mphi ** 2, self.rho_cut,
                           (C.c_double * 4)(*[float(c) for c in self.g0]))

    def x(self, rho, T) -> np.ndarray:
        rho, T = np.broadcast_arrays(np.asarray(rho, dtype=np.float64),
                                     np.asarray(T, dtype=np.float64))
        f1 = np.arcsinh(rho / self.rho_scale) / np.arcsinh(self.rho_max / self.rho_scale)
        f2 = (T - 0.5 * (self.T_lo + self.T_hi)) / (0.5 * (self.T_hi - self.T_lo))
        return np.stack([f1, f2], axis=-1)

    def envelope(self, rho) -> np.ndarray:
        return 1.0 / (1.0 + (np.asarray(rho, dtype=np.float64) / self.rho_cut) ** 2)

    def g_full(self, rho, T) -> np.ndarray:
        """g = (V'/rho - (lam rho^2 - m^2)) / (out_scale e(rho)) in float64."""
        rho = np.asarray(rho, dtype=np.float64)
        vp = P.vprime_full(rho, T, self.model, include_cw=self.include_cw)
        tree = self.model.lam * rho ** 2 - self.model.mphi ** 2
        return (vp / rho - tree) / (self.out_scale * self.envelope(rho))

    def g_target(self, rho, T) -> np.ndarray:
        """The network's target (g - g0(f2)) / f1 in float64."""
        f1 = self.x(rho, T)[..., 0]
        return (self.g_full(rho, T) - self.g0_of(T)) / f1

    def vprime_from_net(self, rho, T, y) -> np.ndarray:
        """V' assembled from network outputs y (float64), as the kernels assemble it."""
        rho = np.asarray(rho, dtype=np.float64)
        f1 = self.x(rho, T)[..., 0]
        b = self.g0_of(T) + f1 * np.asarray(y, dtype=np.float64)
        return rho * (self.model.lam * rho ** 2 - self.model.mphi ** 2
                      + self.out_scale * b * self.envelope(rho))

    def meff2(self, T) -> np.ndarray:
        """The effective mass V'/rho at rho -> 0 of the float64 potential, in GeV^2."""
        T = np.asarray(T, dtype=np.float64)
        return P.vprime_full(np.full_like(T, self.RHO_ZERO), T, self.model,
                             include_cw=self.include_cw) / self.RHO_ZERO

    def vprime_truth(self, rho, T) -> np.ndarray:
        return P.vprime_full(rho, T, self.model, include_cw=self.include_cw)

    def floor(self, T) -> np.ndarray:
        n = max(abs(self.model.coeff_b), abs(self.model.coeff_f), 1.0)
        return 1e-3 * np.asarray(T) ** 4 / (2.0 * np.pi ** 2) * n

    def weights(self, rho, T, barrier_tol: float = 1e-3, floored_tol: float = 1e-4):
        """Per-sample loss weights 1/tolerance^2 on the network output: an error dy moves
        V' by rho m^2 e(rho) f1 dy, which the floored metric divides by max(|V'|, floor)
        (tolerance floored_tol) and the barrier metric (V'/rho in units of m^2, which is what
        the spinodal and the barrier near T_c depend on) by rho m^2 (tolerance barrier_tol).
        Normalised to mean 1."""
        rho = np.asarray(rho, dtype=np.float64)
        T = np.asarray(T, dtype=np.float64)
        k = self.envelope(rho) * self.x(rho, T)[..., 0]
        den = np.maximum(np.abs(self.vprime_truth(rho, T)), self.floor(T))
        k_floor = rho * self.out_scale * k / den
        w = np.maximum((k_floor / floored_tol) ** 2, (k / barrier_tol) ** 2)
        return w / w.mean()

    def sample(self, n: int, seed: int = 0, T_margin: float = 0.0):
        """n (rho, T) samples: half uniform in rho/phi_true, half uniform in f1; T uniform in
        the range widened by T_margin of its width on each side (training uses a margin so the
        fit does not degrade at the range edges).  rho is rounded to float32, the values a
        float32 lattice can hold."""
        rng = np.random.default_rng(seed)
        n1 = n // 2
        rho_a = rng.uniform(0.0, self.rho_max, n1)
        f1 = rng.uniform(0.0, 1.0, n - n1)
        rho_b = self.rho_scale * np.sinh(f1 * np.arcsinh(self.rho_max / self.rho_scale))
        rho = np.concatenate([rho_a, rho_b])
        rho = np.maximum(rho, 1e-3 * self.rho_scale)  # V'/rho is finite at 0; keep rho > 0
        rho = rho.astype(np.float32).astype(np.float64)
        w = (self.T_hi - self.T_lo) * T_margin
        T = rng.uniform(self.T_lo - w, self.T_hi + w, n)
        perm = rng.permutation(n)
        return rho[perm], T[perm]


def init_params(seed: int = 0) -> np.ndarray:
    """Glorot-uniform weights, zero biases (float32)."""
    rng = np.random.default_rng(seed)
    l1, l2, l3 = math.sqrt(6.0 / (2 + H)), math.sqrt(6.0 / (2 * H)), math.sqrt(6.0 / (H + 1))
    return np.concatenate([rng.uniform(-l1, l1, 2 * H), np.zeros(H), rng.uniform(-l2, l2, H * H),
                           np.zeros(H), rng.uniform(-l3, l3, H), [0.0]]).astype(np.float32)


class Surrogate:
    def __init__(self, spec: SurrogateSpec, params: Optional[np.ndarray] = None, seed: int = 0):
        self.spec = spec
        p = init_params(seed) if params is None else np.asarray(params, np.float32)
        if p.size != NPARAMS:
            raise ValueError(f"expected {NPARAMS} parameters, got {p.size}")
        self.params = DeviceArray.from_numpy(p)
        self.m = DeviceArray.zeros((NPARAMS,), np.float32)
        self.v = DeviceArray.zeros((NPARAMS,), np.float32)
        self.t = 0
        self._fm = spec.features()

    # -- training ------------------------------------------------------------------------
    def train(self, rho, T, epochs: int = 200, batch: int = 4096, lr: float = 3e-3,
              lr_end: float = 1e-5, seed: int = 0, weighted: bool = True,
              log=None) -> Dict[str, list]:
        X = self.spec.x(rho, T).astype(np.float32)
        Y = self.spec.g_target(rho, T).astype(np.float32)
        Wt = (self.spec.weights(rho, T) if weighted else np.ones(len(Y))).astype(np.float32)
        n = X.shape[0]
        nb = n // batch
        dX = DeviceArray((n, 2), np.float32)
        dY = DeviceArray((n,), np.float32)
        dW = DeviceArray((n,), np.float32)
        loss = DeviceArray((1,), np.float32)
        rng = np.random.default_rng(seed)
        hist = {"epoch": [], "loss": [], "lr": []}
        total = epochs * nb
        for ep in range(epochs):
            perm = rng.permutation(n)
            dX.upload(X[perm])
            dY.upload(Y[perm])
            dW.upload(Wt[perm])
            for b in range(nb):
                step = ep * nb + b
                cur = lr_end + 0.5 * (lr - lr_end) * (1.0 + math.cos(math.pi * step / total))
                self.t += 1
                _lib.call("tip_mlp_adam_step", self.params.ptr, self.m.ptr, self.v.ptr,
                          dX.ptr + b * batch * 8, dY.ptr + b * batch * 4, dW.ptr + b * batch * 4,
                          batch, cur, 0.9, 0.999, 1e-8, self.t,
                          loss.ptr if b == nb - 1 else None, None)
            l = float(loss.download()[0])
            hist["epoch"].append(ep)
            hist["loss"].append(l)
            hist["lr"].append(cur)
            if log and (ep % max(1, epochs // 10) == 0 or ep == epochs - 1):
                log(f"epoch {ep}: loss {l:.3e} lr {cur:.2e}")
        dX.free()
        dY.free()
        dW.free()
        return hist

    def grad(self, X, Y, W=None):
        X = np.ascontiguousarray(X, np.float32)
        Y = np.ascontiguousarray(Y, np.float32)
        dX, dY = DeviceArray.from_numpy(X), DeviceArray.from_numpy(Y)
        dW = DeviceArray.from_numpy(np.ascontiguousarray(W, np.float32)) if W is not None else None
        g = DeviceArray((NPARAMS,), np.float32)
        loss = DeviceArray((1,), np.float32)
        _lib.call("tip_mlp_grad", self.params.ptr, dX.ptr, dY.ptr,
                  dW.ptr if dW is not None else None, X.shape[0], g.ptr, loss.ptr, None)
        return float(loss.download()[0]), g.download()

    # -- evaluation ----------------------------------------------------------------------
    def forward(self, X) -> np.ndarray:
        X = np.ascontiguousarray(X, np.float32)
        dX = DeviceArray.from_numpy(X)
        dY = DeviceArray((X.shape[0],), np.float32)
        _lib.call("tip_mlp_forward", self.params.ptr, dX.ptr, dY.ptr, X.shape[0], None)
        return dY.download()

    def vprime(self, rho, T) -> np.ndarray:
        rho = np.ascontiguousarray(rho, np.float32).ravel()
        T = np.ascontiguousarray(np.broadcast_to(T, rho.shape), np.float32)
        dr, dT = DeviceArray.from_numpy(rho), DeviceArray.from_numpy(T)
        out = DeviceArray(rho.shape, np.float32)
        _lib.call("tip_mlp_vprime", self.params.ptr, C.byref(self._fm), dr.ptr, dT.ptr, out.ptr,
                  rho.size, None)
        return out.download()

    def tabulate(self, T: float, x_min: float, x_max: float, npts: int, out: DeviceArray = None):
        """Device re-tabulation of V'(x, T) on a uniform grid (float64, odd in x)."""
        out = out if out is not None else DeviceArray((npts,), np.float64)
        _lib.call("tip_mlp_tabulate", self.params.ptr, C.byref(self._fm), float(T), float(x_min),
                  float(x_max), int(npts), out.ptr, None)
        return out

    def numpy_params(self) -> np.ndarray:
        return self.params.download()

    def save(self, path, **meta) -> None:
        sp = self.spec
        np.savez(path, params=self.numpy_params(), rho_scale=sp.rho_scale, rho_cut=sp.rho_cut,
                 rho_frac=sp.rho_frac, T_lo=sp.T_lo, T_hi=sp.T_hi, T_ref=sp.T_ref,
                 model=json.dumps(dataclasses.asdict(sp.model)), include_cw=sp.include_cw,
                 form=FORM, g0=sp.g0, meta=json.dumps(meta))

    @classmethod
    def load(cls, path=None) -> "Surrogate":
        """A saved network (default: the shipped weights for the 6500-7500 K cooling runs)."""
        spec, params = load_params_npz(path or TRAINED_COOLING)
        return cls(spec, params=params)

    def error_report(self, n: int = 1 << 16, seed: int = 99) -> Dict[str, float]:
        """Floored relative error of V' against the float64 truth on fresh samples (float32
        representable rho, T in the nominal range), overall and in the thermal-barrier region
        (rho < 1e5 GeV) relative to rho m^2."""
        rho, T = self.spec.sample(n, seed)
        truth = self.spec.vprime_truth(rho, T)
        got = self.vprime(rho, T).astype(np.float64)
        den = np.maximum(np.abs(truth), self.spec.floor(T))
        rel = np.abs(got - truth) / den
        low = rho < 1e5
        # barrier region on its own scale: |dV'| / (rho * m^2)
        rel_low = np.abs(got - truth)[low] / (rho[low] * self.spec.model.mphi ** 2)
        return {"max_rel_floored": float(rel.max()), "p99_rel_floored": float(np.quantile(rel, 0.99)),
                "median_rel_floored": float(np.median(rel)),
                "max_rel_barrier_region": float(rel_low.max()) if low.any() else 0.0,
                "n": int(n)}


# Weights trained on an MI355X by train_cooling_surrogate(seed=0) (provenance and the error report at
# training time are stored in the file's "meta" entry).
TRAINED_COOLING = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data",
                               "surrogate_cooling_6500_7500.npz")


# Output form of the saved networks: V'/rho = tree + out_scale e (g0 + f1 y)
FORM = "g0-cubic+f1-residual"


def load_params_npz(path: str):
    """(SurrogateSpec, float32 parameters) of a saved network, without touching the GPU."""
    d = np.load(path, allow_pickle=False)
    if "form" not in d.files or str(d["form"]) != FORM:
        raise ValueError(f"{path}: network saved in another output form (expected {FORM})")
    model = P.ModelParams(**json.loads(str(d["model"])))
    spec = SurrogateSpec(model=model, T_lo=float(d["T_lo"]), T_hi=float(d["T_hi"]),
                         T_ref=float(d["T_ref"]), rho_frac=float(d["rho_frac"]),
                         rho_scale=float(d["rho_scale"]), rho_cut=float(d["rho_cut"]),
                         include_cw=bool(d["include_cw"]))
    spec.g0 = np.asarray(d["g0"], np.float64)  # the fit the network was trained against
    return spec, np.asarray(d["params"], np.float32)


def train_cooling_surrogate(epochs: int = 800, n_samples: int = 1 << 20, seed: int = 0, log=None):
    """The recipe for the 7350 -> 6770 K cooling runs: complex-field potential, T in
    [6500, 7500], 2^20 float64-truth samples (5 % T margin), weighted loss (floored 1e-4,
    barrier 1e-3), GPU Adam 800 epochs, batch 1024, cosine lr 2e-3 -> 1e-6."""
    spec = SurrogateSpec()
    rho, T = spec.sample(n_samples, seed=seed, T_margin=0.05)
    sur = Surrogate(spec, seed=seed)
    hist = sur.train(rho, T, epochs=epochs, batch=1024, lr=2e-3, lr_end=1e-6, log=log)
    return sur, hist


def table_error(spec: SurrogateSpec, table, tmin, dinv, T: float, n: int = 1 << 14,
                seed: int = 7) -> Dict[str, float]:
    """The same metrics for a linear-interpolation V' table (what the surrogate replaces)."""
    rho, _ = spec.sample(n, seed)
    Tv = np.full_like(rho, T)
    truth = spec.vprime_truth(rho, Tv)
    t = np.asarray(table, np.float64)
    fidx = (rho - tmin) * dinv
    idx = np.clip(np.trunc(fidx).astype(np.int64), 0, t.size - 2)
    frac = fidx - idx
    got = t[idx] + frac * (t[idx + 1] - t[idx])
    den = np.maximum(np.abs(truth), spec.floor(Tv))
    rel = np.abs(got - truth) / den
    low = rho < 1e5
    rel_low = np.abs(got - truth)[low] / (rho[low] * spec.model.mphi ** 2)
    return {"max_rel_floored": float(rel.max()), "p99_rel_floored": float(np.quantile(rel, 0.99)),
            "median_rel_floored": float(np.median(rel)),
            "max_rel_barrier_region": float(rel_low.max()) if low.any() else 0.0}
