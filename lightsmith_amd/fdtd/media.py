"""ADE media for the FDTD engine (media.py of the reference, restated).

Normalised units: nm, c = eps0 = mu0 = 1 (time in nm/c, omega in rad per nm/c). A medium is
eps(w) = eps_inf + sum_k sigma_k / (w0_k^2 - w^2 - i g_k w); each pole obeys P'' + g P' + w0^2 P = sigma E
(media.py:1-10). A Drude term is the pole with w0 = 0, sigma = wp^2.
"""

from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

HBARC_EV_NM = 197.3269804  # media.py:22


@dataclass(frozen=True)
class Pole:
    """media.py:25-29"""

    sigma: float
    omega0: float
    gamma: float


@dataclass(frozen=True)
class Medium:
    """media.py:32-51"""

    eps_inf: float
    poles: tuple = ()
    label: str = ""

    @property
    def dispersive(self) -> bool:
        return bool(self.poles)

    def eps(self, omega) -> np.ndarray:
        w = np.asarray(omega, dtype=float)
        e = self.eps_inf + 0j * w
        for p in self.poles:
            e = e + p.sigma / (p.omega0 ** 2 - w ** 2 - 1j * p.gamma * w)
        return e

    def describe(self) -> dict:
        return {"label": self.label, "eps_inf": self.eps_inf,
                "poles": [{"sigma": p.sigma, "omega0": p.omega0, "gamma": p.gamma} for p in self.poles]}


def drude_lorentz(eps_inf: float, wp_ev: float = 0.0, gamma_ev: float = 0.0, lorentz=(), label: str = "") -> Medium:
    """A Drude term (wp, gamma in eV) plus Lorentz terms (delta_eps, omega0_ev, gamma_ev) as poles
    (media.py:68-75)."""
    poles = []
    if wp_ev > 0:
        poles.append(Pole((wp_ev / HBARC_EV_NM) ** 2, 0.0, gamma_ev / HBARC_EV_NM))
    for de, w0_ev, g_ev in lorentz:
        w0 = w0_ev / HBARC_EV_NM
        poles.append(Pole(de * w0 ** 2, w0, g_ev / HBARC_EV_NM))
    return Medium(float(eps_inf), tuple(poles), label)


def pole_coefficients(pole: Pole, dt: float) -> tuple[float, float, float]:
    """(c1, c2, c3) of P^{n+1} = c1 P^n + c2 P^{n-1} + c3 E^n, central differences (engine.py:108-111)."""
    den = 1 + pole.gamma * dt / 2
    return ((2 - pole.omega0 ** 2 * dt ** 2) / den, -(1 - pole.gamma * dt / 2) / den, pole.sigma * dt ** 2 / den)


def discrete_eps(m: Medium, omega, dt: float) -> np.ndarray:
    """The permittivity the D-form ADE update realises at omega (layered.py:24-31)."""
    Z = np.exp(-1j * np.asarray(omega, dtype=float) * dt)
    eps = m.eps_inf + 0j * Z
    for p in m.poles:
        c1, c2, c3 = pole_coefficients(p, dt)
        eps = eps + c3 / (Z - c1 - c2 / Z)
    return eps


def _amplification(m: Medium, dt: float, kappa: np.ndarray) -> float:
    """Spectral radius of one Yee + D-form ADE step over curl-curl eigenvalues kappa (setup.py:26-47)."""
    k = len(m.poles)
    n = 2 + 2 * k
    e = np.zeros(n)
    e[0] = 1 / m.eps_inf
    e[2::2] = -1 / m.eps_inf
    A = np.zeros((kappa.size, n, n))
    A[:, 0] = -(dt ** 2) * kappa[:, None] * e
    A[:, 0, 0] += 2
    A[:, 0, 1] -= 1
    A[:, 1, 0] = 1
    for j, p in enumerate(m.poles):
        c1, c2, c3 = pole_coefficients(p, dt)
        i = 2 + 2 * j
        A[:, i] = c3 * e
        A[:, i, i] += c1
        A[:, i, i + 1] += c2
        A[:, i + 1, i] = 1
    return float(np.abs(np.linalg.eigvals(A)).max())


def dispersive_dt_limit(media, kappa_max: float) -> float:
    """Largest stable dt of the explicit ADE update over the dispersive media (setup.py:50-69)."""
    kappa = np.linspace(0.0, kappa_max, 65)
    limit = math.inf
    for m in media:
        if not m.dispersive:
            continue
        hi = 2 / math.sqrt(kappa_max / m.eps_inf)
        if _amplification(m, hi, kappa) <= 1 + 1e-6:
            continue
        lo = 0.0
        for _ in range(40):
            mid = 0.5 * (lo + hi)
            lo, hi = (mid, hi) if _amplification(m, mid, kappa) <= 1 + 1e-6 else (lo, mid)
        limit = min(limit, lo)
    return limit


def limit_dt(dt: float, courant: float, media, kappa_max: float) -> float:
    """setup.py:72-75"""
    limit = dispersive_dt_limit(media, kappa_max)
    return dt if dt <= limit else courant * limit


def time_step(dx: float, courant: float, media) -> float:
    """Courant step of a uniform cubic grid, lowered for unstable ADE media (setup.py:78-91)."""
    eps_min = min([1.0] + [m.eps_inf for m in media])
    return limit_dt(courant * dx * math.sqrt(eps_min) / math.sqrt(3), courant, media, 12 / dx ** 2)
