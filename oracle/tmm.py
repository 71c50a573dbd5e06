"""Coherent multilayer reflectance and transmittance (TEST INFRASTRUCTURE ONLY: the whole-solve check of the layer-stack
and periodic drivers, which have no other closed form).

The reference's TMM backend (backends/tmm.py:64-90) climbs the stack interface by interface with the Airy sum;
this restates the same physics by the characteristic-matrix method of Born & Wolf, "Principles of Optics",
section 1.6 (Macleod, "Thin-Film Optical Filters", chapter 2): with q_j = sqrt(eps_j - (n_0 sin theta_0)^2)
(Im q_j >= 0, decaying along +z for the e^{-i omega t} time convention), the tangential-field admittance
eta_j = q_j (s) or eps_j / q_j (p) and the phase thickness delta_j = k0 q_j d_j, each layer maps the tangential
(E, H) at its far face to its near face by [[cos delta, -i sin delta / eta], [-i eta sin delta, cos delta]]; with
(B, C) = prod_j M_j (1, eta_sub):
    r = (eta_0 B - C) / (eta_0 B + C),  t = 2 eta_0 / (eta_0 B + C),
    R = |r|^2,  T = Re(eta_sub) / Re(eta_0) |t|^2.

Pinned in tests/test_oracle_pins.py against the three-layer Airy sum of tests/test_drivers.py (normal and oblique,
s and p, lossless and Drude films), the quarter-wave mirror closed form R = ((n0 - Y) / (n0 + Y))^2 with
Y = (n_H / n_L)^{2N} n_s, and energy conservation of lossless stacks.
"""

from __future__ import annotations

import numpy as np


def _q(eps, s2):
    q = np.sqrt(np.asarray(eps, dtype=complex) - s2)
    return np.where(q.imag < 0, -q, q)


def stack_rt(eps_layers, thick_nm, eps_sup, eps_sub, wl_nm, theta_rad: float, pol: str):
    """(R, T) per wavelength of layers (permittivity per layer per wavelength, thickness nm) between a lossless
    superstrate (light comes from it) and a substrate, incidence angle theta in the superstrate."""
    wl = np.atleast_1d(np.asarray(wl_nm, dtype=float))
    k0 = 2 * np.pi / wl
    eps0 = np.broadcast_to(np.asarray(eps_sup, dtype=complex), wl.shape)
    epss = np.broadcast_to(np.asarray(eps_sub, dtype=complex), wl.shape)
    s2 = eps0.real * np.sin(theta_rad) ** 2

    def eta(eps):
        q = _q(eps, s2)
        return q if pol == "s" else eps / q

    B = np.ones(wl.shape, dtype=complex)
    C = eta(epss).astype(complex)
    for eps_j, d in zip(reversed(list(eps_layers)), reversed(list(thick_nm))):
        e = np.broadcast_to(np.asarray(eps_j, dtype=complex), wl.shape)
        q = _q(e, s2)
        y = q if pol == "s" else e / q
        delta = k0 * q * float(d)
        c, sn = np.cos(delta), np.sin(delta)
        B, C = c * B - 1j * sn / y * C, -1j * y * sn * B + c * C
    y0 = eta(eps0)
    den = y0 * B + C
    r = (y0 * B - C) / den
    t = 2 * y0 / den
    R = np.abs(r) ** 2
    T = np.real(eta(epss)) / np.real(y0) * np.abs(t) ** 2
    return R, T
