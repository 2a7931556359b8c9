"""Synthetic inputs and the benchmark configurations of this hot path.

Fields: a seeded Gaussian random field with isotropic k^-3 spectrum that is
zero on the boundary, built from i.i.d. N(0, 1) DST-I coefficients c_pq
scaled by K^-2 (K^2 = p^2 + q^2) and inverted with a DST-I; each layer is
rescaled to the requested RMS.  The generator is numpy's counter-based
Philox, seeded with 0x5EED0000 + config_id * 1000 + member.

Non-dimensionalisation follows the reference's driver (qg3.F90:150-168,
236): length scale = grid spacing dx, velocity scale 1 m/s, time scale
dx / (1 m/s); CGS inputs of the reference map to the SI values here.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from . import strat as _strat


def seed_for(config_id: int, member: int = 0) -> int:
    return 0x5EED0000 + config_id * 1000 + member


def _dst1(a: np.ndarray, axis: int) -> np.ndarray:
    """Unnormalised DST-I along one axis via a real FFT of the odd extension."""
    a = np.moveaxis(a, axis, -1)
    n = a.shape[-1]
    z = np.zeros(a.shape[:-1] + (2 * (n + 1),))
    z[..., 1:n + 1] = a
    z[..., n + 2:] = -a[..., ::-1]
    out = -np.fft.rfft(z, axis=-1).imag[..., 1:n + 1]
    return np.moveaxis(out, -1, axis)


def k3_field(n: int, layers: int, seed: int, rms: float = 1e-5) -> np.ndarray:
    """(n+2, n+2, L) Fortran-ordered field, zero on the boundary, with an
    isotropic k^-3 spectrum and the given RMS per layer (over the interior)."""
    rng = np.random.Generator(np.random.Philox(seed))
    p = np.arange(1, n + 1, dtype=np.float64)
    K2 = p[:, None] ** 2 + p[None, :] ** 2
    out = np.zeros((n + 2, n + 2, layers), order="F")
    for l in range(layers):
        c = rng.standard_normal((n, n)) / K2
        f = _dst1(_dst1(c, 0), 1)
        f *= rms / np.sqrt(np.mean(f * f))
        out[1:-1, 1:-1, l] = f
    return out


@dataclass
class QGConfig:
    name: str
    n: int                      # interior points per side (DST length)
    dx: float                   # grid spacing [m]
    H: list                     # layer depths [m]
    gprime: list                # reduced gravities g'_{i+1/2} [m s^-2]
    f0: float = 1e-4
    beta: float = 2e-11         # [m^-1 s^-1]
    dt: float = 1800.0          # [s]
    tau0: float = 0.05          # wind stress [N m^-2]
    rho0: float = 1000.0        # [kg m^-3] (assumed; not stated in the configs)
    drag: float = 1e-7          # bottom drag [s^-1]
    visc: float = 100.0         # lateral viscosity [m^2 s^-1] (reference default)
    alpha: float = 120e3        # partial-slip length [m] (reference default)
    rms: float = 1e-5           # PV RMS [s^-1]
    config_id: int = 0
    extra: dict = field(default_factory=dict)

    @property
    def g(self) -> int:
        return self.n + 2

    @property
    def layers(self) -> int:
        return len(self.H)

    def stratification(self) -> _strat.Stratification:
        return _strat.from_density(self.H, self.gprime, self.f0, self.dx)

    # non-dimensional parameters (qg3.F90:150-168)
    @property
    def t_scale(self) -> float:
        return self.dx / 1.0

    def nondim(self) -> dict:
        st = self.stratification()
        h1 = self.H[0] / self.dx
        tau = (self.tau0 / self.rho0) / (1.0 * 1.0 * (self.n + 1)) / h1
        return dict(
            dt=self.dt / self.t_scale,
            beta=np.full(self.layers, self.beta * self.dx * self.dx / 1.0),
            viscosity=self.visc / (self.dx * 1.0),
            visc_bot=self.drag * self.dx / 1.0,
            alpha=self.dx / self.alpha,
            tau=tau,
            strat=st,
        )

    def wind(self, shift=0.0, tilt=0.0, asym=1.0) -> np.ndarray:
        """force_wind(G, G) (qg3.F90:253-279)."""
        g = self.g
        tau = self.nondim()["tau"]
        a_w, b_w = -1.0 * asym, 1.0 / asym
        i = np.arange(g)[:, None]
        j = np.arange(g)[None, :]
        x = j / (g - 1) - 0.5
        y = -i / (g - 1) + 0.5
        ym = shift + tilt * x
        f = np.where(y < ym, a_w * np.sin(np.pi * (0.5 + y) / (0.5 + ym)),
                     b_w * np.sin(np.pi * (y - ym) / (0.5 - ym)))
        return np.asfortranarray(tau * 2.0 * np.pi * f)

    def q(self, member: int = 0) -> np.ndarray:
        """Seeded PV anomaly, non-dimensional (q * t_scale)."""
        return k3_field(self.n, self.layers, seed_for(self.config_id, member), self.rms) * self.t_scale


def config(name: str) -> QGConfig:
    """The hot path's workloads by name: "qg3_255" (3-layer, 255^2 interior,
    dx 10 km; CPU-runnable), "qg3_1023" (3-layer, 1023^2, dx 5 km) and
    "qg6_2047" (6-layer, 2047^2, dx 2.5 km)."""
    if name == "qg3_255":
        return QGConfig("3-layer 255^2, dx 10 km", 255, 10e3, [300.0, 700.0, 3000.0],
                        [0.02, 0.01], config_id=0)
    if name == "qg3_1023":
        return QGConfig("3-layer 1023^2, dx 5 km", 1023, 5e3, [300.0, 700.0, 3000.0],
                        [0.02, 0.01], config_id=1)
    if name == "qg6_2047":
        return QGConfig("6-layer 2047^2, dx 2.5 km", 2047, 2.5e3,
                        [150.0, 250.0, 400.0, 700.0, 1200.0, 2300.0],
                        list(np.linspace(0.03, 0.005, 5)), config_id=2)
    raise KeyError(name)


def steep_member(layers: int, ratio: float, n: int = 127, member: int = 99):
    """Member whose layer depths grow by `ratio` per layer (summing to
    4000 m; g' linear 0.03 -> 0.005): an ill-conditioned layer <-> mode
    projection for large ratios (8 layers, ratio 2.5: kappa_inf 85.8)."""
    H = np.array([ratio ** k for k in range(layers)])
    H = list(4000.0 * H / H.sum())
    gp = list(np.linspace(0.03, 0.005, layers - 1))
    return QGConfig(f"steep {layers}-layer member", n, 5e3, H, gp, rms=1e-5, config_id=3,
                    extra={"member": member})


def ensemble_member(member: int, seed_base: int = 3):
    """Member of the 64-member mixed ensemble: interior size from
    {255, 511, 767, 1023}, 2..8 layers, H geometric summing to 4000 m,
    g' linear 0.03 -> 0.005, dx 5 km, PV amplitude 1e-5 s^-1 +-10 %."""
    rng = np.random.Generator(np.random.Philox(seed_for(seed_base, member)))
    n = int(rng.choice([255, 511, 767, 1023]))
    L = int(rng.integers(2, 9))
    r = 1.6
    H = np.array([r ** k for k in range(L)])
    H = list(4000.0 * H / H.sum())
    gp = list(np.linspace(0.03, 0.005, L - 1))
    amp = float(rng.uniform(0.9, 1.1))
    return QGConfig(f"ensemble member {member}", n, 5e3, H, gp, rms=1e-5 * amp,
                    config_id=seed_base, extra={"member": member})
