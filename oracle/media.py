"""ADE media and time-step rules, restated in FP64 (oracle; test infrastructure only).

Normalised units: nm, c = eps0 = mu0 = 1, time in nm/c, omega in rad per nm/c.
A pole obeys P'' + g P' + w0^2 P = sigma E, so eps(w) = eps_inf + sum sigma / (w0^2 - w^2 - i g w)
(media.py:1-10).
"""

from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

HBARC_EV_NM = 197.3269804  # media.py:22


@dataclass(frozen=True)
class Pole:  # media.py:25-29
    sigma: float
    omega0: float
    gamma: float


@dataclass(frozen=True)
class Medium:  # media.py:32-51
    eps_inf: float
    poles: tuple = ()

    @property
    def dispersive(self) -> bool:
        return len(self.poles) > 0

    def eps(self, omega):
        w = np.asarray(omega, dtype=float)
        out = self.eps_inf + 0j * w
        for p in self.poles:
            out = out + p.sigma / (p.omega0 ** 2 - w ** 2 - 1j * p.gamma * w)
        return out


def drude_lorentz(eps_inf: float, wp_ev: float = 0.0, gamma_ev: float = 0.0, lorentz=()) -> Medium:
    """Drude term plus Lorentz terms (delta_eps, omega0_ev, gamma_ev) as poles (media.py:68-75)."""
    poles = []
    if wp_ev > 0:
        poles.append(Pole((wp_ev / HBARC_EV_NM) ** 2, 0.0, gamma_ev / HBARC_EV_NM))
    for de, w0_ev, g_ev in lorentz:
        w0 = w0_ev / HBARC_EV_NM
        poles.append(Pole(de * w0 ** 2, w0, g_ev / HBARC_EV_NM))
    return Medium(float(eps_inf), tuple(poles))


def pole_coefficients(p: Pole, dt: float):
    """P^{n+1} = c1 P^n + c2 P^{n-1} + c3 E^n by central differences (engine.py:108-111)."""
    den = 1.0 + p.gamma * dt / 2.0
    return ((2.0 - p.omega0 ** 2 * dt ** 2) / den, -(1.0 - p.gamma * dt / 2.0) / den, p.sigma * dt ** 2 / den)


def discrete_eps(m: Medium, omega, dt: float):
    """eps of the D-form ADE update at omega: eps_inf + sum c3 / (Z - c1 - c2 / Z), Z = exp(-i omega dt)
    (layered.py:24-31)."""
    Z = np.exp(-1j * np.asarray(omega, dtype=float) * dt)
    out = m.eps_inf + 0j * Z
    for p in m.poles:
        c1, c2, c3 = pole_coefficients(p, dt)
        out = out + c3 / (Z - c1 - c2 / Z)
    return out


def amplification(m: Medium, dt: float, kappa) -> float:
    """Spectral radius of one Yee + D-form ADE step over curl-curl eigenvalues kappa (setup.py:26-47).

    State (D^n, D^{n-1}, P_1^n, P_1^{n-1}, ...); D^{n+1} = 2 D^n - D^{n-1} - dt^2 kappa E^n with
    E^n = (D^n - sum_k P_k^n) / eps_inf."""
    kappa = np.atleast_1d(np.asarray(kappa, dtype=float))
    npole = len(m.poles)
    size = 2 + 2 * npole
    e_row = np.zeros(size)
    e_row[0] = 1.0 / m.eps_inf
    e_row[2::2] = -1.0 / m.eps_inf
    A = np.zeros((kappa.size, size, size))
    A[:, 0, :] = -(dt ** 2) * kappa[:, None] * e_row[None, :]
    A[:, 0, 0] += 2.0
    A[:, 0, 1] -= 1.0
    A[:, 1, 0] = 1.0
    for j, p in enumerate(m.poles):
        c1, c2, c3 = pole_coefficients(p, dt)
        r = 2 + 2 * j
        A[:, r, :] = c3 * e_row[None, :]
        A[:, r, r] += c1
        A[:, r, r + 1] += c2
        A[:, r + 1, r] = 1.0
    return float(np.abs(np.linalg.eigvals(A)).max())


def dispersive_dt_limit(media, kappa_max: float) -> float:
    """Largest stable dt over the dispersive media by bisection (setup.py:50-69)."""
    kappa = np.linspace(0.0, kappa_max, 65)
    limit = math.inf
    for m in media:
        if not m.dispersive:
            continue
        hi = 2.0 / math.sqrt(kappa_max / m.eps_inf)
        if amplification(m, hi, kappa) <= 1 + 1e-6:
            continue
        lo = 0.0
        for _ in range(40):
            mid = 0.5 * (lo + hi)
            if amplification(m, mid, kappa) <= 1 + 1e-6:
                lo = mid
            else:
                hi = mid
        limit = min(limit, lo)
    return limit


def time_step(dx: float, courant: float, media) -> float:
    """Courant step on a uniform cubic grid, lowered for unstable ADE media (setup.py:72-91)."""
    eps_min = min([1.0] + [m.eps_inf for m in media])
    dt = courant * dx * math.sqrt(eps_min) / math.sqrt(3.0)
    limit = dispersive_dt_limit(media, 12.0 / dx ** 2)
    return dt if dt <= limit else courant * limit
