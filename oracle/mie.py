"""Mie cross sections of a homogeneous sphere (TEST INFRASTRUCTURE ONLY: the whole-solve check of the scattering
driver).

The reference's own Mie (backends/mie_native.py:216-262) cannot run here (array_api_compat is missing), so this
restates the published series of Bohren & Huffman, "Absorption and Scattering of Light by Small Particles" (1983),
chapter 4 and appendix A (BHMIE): the logarithmic derivative D_n(mx) by downward recurrence from n_stop + 15, the
Riccati-Bessel functions psi_n, xi_n by upward recurrence, a_n = ((D_n/m + n/x) psi_n - psi_{n-1}) /
((D_n/m + n/x) xi_n - xi_{n-1}), b_n likewise with m D_n, and
Q_ext = 2/x^2 sum (2n+1) Re(a_n + b_n), Q_sca = 2/x^2 sum (2n+1)(|a_n|^2 + |b_n|^2), n_stop = x + 4 x^(1/3) + 2.

Pinned in tests/test_oracle_pins.py by the book's worked example (m = 1.55, x = 5.213), the Rayleigh limit and the
optical theorem for a lossless sphere.
"""

from __future__ import annotations

import numpy as np


def mie_ab(m: complex, x: float, n_max: int) -> tuple[np.ndarray, np.ndarray]:
    """The scattering coefficients a_n, b_n, n = 1 .. n_max (the same recurrences as mie_q)."""
    m, x = complex(m), float(x)
    mx = m * x
    n_mx = int(max(n_max, abs(mx))) + 15
    D = np.zeros(n_mx + 1, dtype=complex)
    for n in range(n_mx, 0, -1):
        D[n - 1] = n / mx - 1.0 / (D[n] + n / mx)
    psi0, psi1 = np.cos(x), np.sin(x)
    chi0, chi1 = -np.sin(x), np.cos(x)
    xi1 = complex(psi1, -chi1)
    a, b = np.zeros(n_max, dtype=complex), np.zeros(n_max, dtype=complex)
    for n in range(1, n_max + 1):
        psi = (2.0 * n - 1.0) * psi1 / x - psi0
        chi = (2.0 * n - 1.0) * chi1 / x - chi0
        xi = complex(psi, -chi)
        a[n - 1] = ((D[n] / m + n / x) * psi - psi1) / ((D[n] / m + n / x) * xi - xi1)
        b[n - 1] = ((m * D[n] + n / x) * psi - psi1) / ((m * D[n] + n / x) * xi - xi1)
        psi0, psi1 = psi1, psi
        chi0, chi1 = chi1, chi
        xi1 = complex(psi1, -chi1)
    return a, b


def mie_q(m: complex, x: float) -> tuple[float, float, float]:
    """(Q_ext, Q_sca, Q_back) for relative index m and size parameter x = k r (k in the host medium)."""
    m = complex(m)
    x = float(x)
    mx = m * x
    n_stop = int(x + 4.0 * x ** (1.0 / 3.0) + 2.0)
    n_mx = int(max(n_stop, abs(mx))) + 15
    D = np.zeros(n_mx + 1, dtype=complex)
    for n in range(n_mx, 0, -1):
        D[n - 1] = n / mx - 1.0 / (D[n] + n / mx)
    psi0, psi1 = np.cos(x), np.sin(x)
    chi0, chi1 = -np.sin(x), np.cos(x)
    xi1 = complex(psi1, -chi1)
    q_ext = q_sca = 0.0
    back = 0j
    for n in range(1, n_stop + 1):
        psi = (2.0 * n - 1.0) * psi1 / x - psi0
        chi = (2.0 * n - 1.0) * chi1 / x - chi0
        xi = complex(psi, -chi)
        a = ((D[n] / m + n / x) * psi - psi1) / ((D[n] / m + n / x) * xi - xi1)
        b = ((m * D[n] + n / x) * psi - psi1) / ((m * D[n] + n / x) * xi - xi1)
        q_ext += (2 * n + 1) * (a + b).real
        q_sca += (2 * n + 1) * (abs(a) ** 2 + abs(b) ** 2)
        back += (2 * n + 1) * (-1) ** n * (a - b)
        psi0, psi1 = psi1, psi
        chi0, chi1 = chi1, chi
        xi1 = complex(psi1, -chi1)
    s = 2.0 / (x * x)
    return s * q_ext, s * q_sca, abs(back) ** 2 / (x * x)


def cross_sections(eps_sphere, radius: float, wavelengths, n_host: float = 1.0):
    """(C_ext, C_sca, C_abs) per wavelength (vacuum, same length unit as the radius) for a sphere of complex
    permittivity eps_sphere (a scalar or one value per wavelength) in a host of index n_host."""
    wl = np.atleast_1d(np.asarray(wavelengths, dtype=float))
    eps = np.broadcast_to(np.asarray(eps_sphere, dtype=complex), wl.shape)
    out = np.zeros((3, len(wl)))
    for q, (lam, e) in enumerate(zip(wl, eps)):
        k = 2 * np.pi * n_host / lam
        qe, qs, _ = mie_q(np.sqrt(e) / n_host, k * radius)
        g = np.pi * radius ** 2
        out[:, q] = (qe * g, qs * g, (qe - qs) * g)
    return out[0], out[1], out[2]


def amplitudes(m: complex, x: float, mu) -> tuple[np.ndarray, np.ndarray]:
    """The scattering amplitudes S1, S2 at mu = cos(theta) (Bohren & Huffman eq. 4.74): with the angular functions
    pi_n = (2n-1)/(n-1) mu pi_{n-1} - n/(n-1) pi_{n-2} (pi_0 = 0, pi_1 = 1) and tau_n = n mu pi_n - (n+1) pi_{n-1},
    S1 = sum (2n+1)/(n(n+1)) (a_n pi_n + b_n tau_n), S2 = sum (2n+1)/(n(n+1)) (a_n tau_n + b_n pi_n)."""
    mu = np.atleast_1d(np.asarray(mu, dtype=float))
    n_stop = int(x + 4.0 * x ** (1.0 / 3.0) + 2.0)
    a, b = mie_ab(m, x, n_stop)
    s1 = np.zeros(mu.shape, dtype=complex)
    s2 = np.zeros(mu.shape, dtype=complex)
    pi0, pi1 = np.zeros_like(mu), np.ones_like(mu)
    for n in range(1, n_stop + 1):
        tau = n * mu * pi1 - (n + 1) * pi0
        f = (2 * n + 1) / (n * (n + 1))
        s1 += f * (a[n - 1] * pi1 + b[n - 1] * tau)
        s2 += f * (a[n - 1] * tau + b[n - 1] * pi1)
        pi0, pi1 = pi1, ((2 * n + 1) * mu * pi1 - (n + 1) * pi0) / n
    return s1, s2


def dcsca_domega(eps_sphere, radius: float, wavelength: float, theta_deg, phi_deg, n_host: float = 1.0):
    """dC_sca/dOmega (n_theta, n_phi) of a sphere, theta from the incident direction, phi from the incident
    polarisation: (|S2|^2 cos^2 phi + |S1|^2 sin^2 phi) / k^2 (Bohren & Huffman eq. 3.16, 4.74)."""
    k = 2 * np.pi * n_host / wavelength
    s1, s2 = amplitudes(np.sqrt(complex(eps_sphere)) / n_host, k * radius, np.cos(np.radians(theta_deg)))
    ph = np.radians(np.asarray(phi_deg, dtype=float))[None, :]
    return (np.abs(s2[:, None]) ** 2 * np.cos(ph) ** 2 + np.abs(s1[:, None]) ** 2 * np.sin(ph) ** 2) / k ** 2


def pole(n: int, eps_fn, radius: float, n_host: float, omega0: complex, tol: float = 1e-13) -> complex:
    """A complex resonance frequency (rad/nm, exp(-i omega t): Im < 0) of Mie's a_n: a zero of its denominator
    m psi_n(m x) xi_n'(x) - xi_n(x) psi_n'(m x), x = n_host omega r, m = sqrt(eps_fn(omega)) / n_host (even in m, so
    the branch does not matter), found by the secant method from omega0."""
    from scipy.special import spherical_jn, spherical_yn

    def denom(w):
        x = n_host * w * radius
        m = np.sqrt(complex(eps_fn(w))) / n_host
        z = m * x
        jz, djz = spherical_jn(n, z), spherical_jn(n, z, True)
        hx = spherical_jn(n, x) + 1j * spherical_yn(n, x)
        dhx = spherical_jn(n, x, True) + 1j * spherical_yn(n, x, True)
        psi_z, dpsi_z = z * jz, jz + z * djz
        xi_x, dxi_x = x * hx, hx + x * dhx
        return m * psi_z * dxi_x - xi_x * dpsi_z

    w0, w1 = complex(omega0), complex(omega0) * (1 + 1e-4)
    f0, f1 = denom(w0), denom(w1)
    for _ in range(200):
        w2 = w1 - f1 * (w1 - w0) / (f1 - f0)
        if abs(w2 - w1) < tol * abs(w2):
            return w2
        w0, f0, w1, f1 = w1, f1, w2, denom(w2)
    raise RuntimeError("the secant iteration for the Mie pole did not converge")

