"""Scattering by a small sphere above a planar substrate in the dipole limit (TEST INFRASTRUCTURE ONLY: the
whole-solve check of the scattering driver over a substrate, the model the reference checks its FDTD against in
test_fdtd_substrate.py:72-130).

A sphere of radius r and index n_p, centred h above the plane z = 0 between a host n1 (above) and a substrate n2,
is a point dipole p = alpha E_loc with the electric-dipole polarizability of Mie theory, alpha = 6 pi i a_1 / k1^3
(Bohren & Huffman eq. 4.56 for a_1; alpha -> 4 pi r^3 (m^2 - 1)/(m^2 + 2) for x -> 0), in units where the dipole's
field is (1/4 pi) times the free dyadic Green's function. It is driven by the background plane wave at its centre
(incident + reflected from above; transmitted when lit through the substrate; Fresnel amplitudes in the
e^{-i omega t} convention) and by its own reflected field, E_loc = E_bg + G_R p. The reflected Green's function at
the dipole, diagonal by symmetry, is the Sommerfeld integral over the in-plane wavenumber s = k_rho / k1
(Novotny & Hecht, "Principles of Nano-Optics", 2nd ed., section 10.4):

    G_par  = i k1^3 / (8 pi) int_0^inf (s / s_z) [r_s(s) - s_z^2 r_p(s)] e^{2 i k1 h s_z} ds,
    G_perp = i k1^3 / (4 pi) int_0^inf (s^3 / s_z) r_p(s) e^{2 i k1 h s_z} ds,

s_z = sqrt(1 - s^2) (Im >= 0), r_s, r_p the Fresnel coefficients of the tangential E (s) and H (p) fields; the
integral is taken over s = sin t (t in [0, pi/2]) and s = cosh t (s_z = i sinh t, out to e^{-40}). So
p_j = alpha E_bg,j / (1 - alpha G_j), and the power it radiates, into both half-spaces, is the free dipole's times
1 + (6 pi / k1^3) Im G_j per component (Novotny & Hecht eq. 10.26). Per unit incident intensity (n_inc |E|^2 / 2):

    C_sca = (n1 / n_inc) k1^4 / (6 pi) sum_j |p_j|^2 (1 + 6 pi Im G_j / k1^3).

Pinned in tests/test_oracle_pins.py: with the substrate equal to the host it is the Mie series' a_1 term; over a
perfect conductor (r_s = -1, r_p = 1) the Sommerfeld integrals are the image dipole's field in closed form; the
Fresnel amplitudes it drives the dipole with reflect |r|^2 = TMM's R.
"""

from __future__ import annotations

import numpy as np
from scipy.integrate import quad

from .mie import mie_ab


def fresnel(e1: complex, e2: complex, s):
    """(r_s, r_p) at normalised in-plane wavenumber s = k_x / k1 for a wave in medium 1 meeting medium 2:
    r_s = (q1 - q2) / (q1 + q2), r_p = (e2 q1 - e1 q2) / (e2 q1 + e1 q2), q = k_z / k1."""
    q1 = np.sqrt(complex(1.0 - s * s))
    q2 = np.sqrt(complex(e2 / e1 - s * s))
    q2 = -q2 if q2.imag < 0 else q2
    return (q1 - q2) / (q1 + q2), (e2 * q1 - e1 * q2) / (e2 * q1 + e1 * q2)


def reflected_green(k1: float, h: float, rs_fn, rp_fn) -> tuple[complex, complex]:
    """(G_par, G_perp) of a dipole h above a plane whose reflection is r_s(s), r_p(s)."""

    def integrand(s, sz):
        rs, rp = rs_fn(s), rp_fn(s)
        ph = np.exp(2j * k1 * h * sz)
        return (s / sz) * (rs - sz * sz * rp) * ph, (s ** 3 / sz) * rp * ph

    def cquad(f, a, b):
        re = quad(lambda t: f(t).real, a, b, limit=600)[0]
        im = quad(lambda t: f(t).imag, a, b, limit=600)[0]
        return re + 1j * im

    t_hi = float(np.arcsinh(20.0 / (k1 * h)))
    total = []
    for i in range(2):
        prop = cquad(lambda t: integrand(np.sin(t), np.cos(t))[i] * np.cos(t), 0.0, np.pi / 2)
        evan = cquad(lambda t: integrand(np.cosh(t), 1j * np.sinh(t))[i] * np.sinh(t), 0.0, t_hi)
        total.append(prop + evan)
    return 1j * k1 ** 3 / (8 * np.pi) * total[0], 1j * k1 ** 3 / (4 * np.pi) * total[1]


def background_field(n1: float, n2: float, k0: float, kx: float, h: float, pol: str, from_below: bool):
    """(E at the dipole (3,), n_inc): the plane wave (unit amplitude in its incidence medium, phase 0 at the
    interface, in the xz plane) with its reflection, or the transmitted wave when lit from below."""
    e1, e2 = n1 * n1, n2 * n2
    k1, k2 = n1 * k0, n2 * k0
    q1 = np.sqrt(complex(k1 * k1 - kx * kx))
    q2 = np.sqrt(complex(k2 * k2 - kx * kx))
    if from_below:
        if pol == "s":
            return np.array([0.0, 2 * q2 / (q2 + q1) * np.exp(1j * q1 * h), 0.0]), n2
        t_h = 2 * e1 * q2 / (e1 * q2 + e2 * q1)  # H-field transmission, substrate to host
        amp = t_h * (e2 / e1) / k2 * np.exp(1j * q1 * h)
        return np.array([q1 * amp, 0.0, -kx * amp]), n2
    down, up = np.exp(-1j * q1 * h), np.exp(1j * q1 * h)
    if pol == "s":
        return np.array([0.0, down + (q1 - q2) / (q1 + q2) * up, 0.0]), n1
    r_h = (e2 * q1 - e1 * q2) / (e2 * q1 + e1 * q2)
    return np.array([-(q1 / k1) * (down - r_h * up), 0.0, -(kx / k1) * (down + r_h * up)]), n1


def sphere_above_substrate(r: float, n_p: float, n1: float, n2: float, h: float, wavelengths, theta_deg: float = 0.0,
                           pol: str = "s", from_below: bool = False) -> np.ndarray:
    """C_sca (nm^2) per wavelength of the sphere centred h above the substrate, lit in the xz plane at theta from
    the normal of its incidence medium (s: E along y; p: in the xz plane)."""
    out = []
    th = np.radians(theta_deg)
    e1, e2 = n1 * n1, n2 * n2
    for lam in np.atleast_1d(np.asarray(wavelengths, dtype=float)):
        k0 = 2 * np.pi / lam
        k1 = n1 * k0
        kx = (n2 if from_below else n1) * k0 * np.sin(th)
        e_bg, n_inc = background_field(n1, n2, k0, kx, h, pol, from_below)
        a1 = mie_ab(n_p / n1, k1 * r, 1)[0][0]
        alpha = 6j * np.pi * a1 / k1 ** 3
        g_par, g_perp = reflected_green(k1, h, lambda s: fresnel(e1, e2, s)[0], lambda s: fresnel(e1, e2, s)[1])
        g = np.array([g_par, g_par, g_perp])
        p = alpha * e_bg / (1 - alpha * g)
        power = np.sum(np.abs(p) ** 2 * (1 + 6 * np.pi / k1 ** 3 * g.imag))
        out.append((n1 / n_inc) * k1 ** 4 / (6 * np.pi) * power)
    return np.array(out)
