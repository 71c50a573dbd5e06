"""Decay rates of a dipole outside a sphere (TEST INFRASTRUCTURE ONLY: the whole-solve check of the emission driver;
the reference checks its FDTD emission against its Mie backend, test_emission.py:70-115).

The reference's closed form (backends/mie_native.py ``dipole_emission``) is not runnable here; this restates the
published series (Chew, J. Chem. Phys. 87, 1355 (1987); Kim, Leung & Liu, Surf. Sci. 195, 1 (1988)) with the
Mie coefficients a_n (electric) and b_n (magnetic) of Bohren & Huffman (eq. 4.53) for the sphere in the host, the
host wavenumber k, the dipole at distance d from the centre, y = k d, h_n = j_n + i y_n, psi_n = y j_n,
xi_n = y h_n (e^{-i omega t}):

    total,     radial:      1 - 3/2 Re sum n(n+1)(2n+1) a_n (h_n(y) / y)^2
    total,     tangential:  1 - 3/4 Re sum (2n+1) [b_n h_n(y)^2 + a_n (xi_n'(y) / y)^2]
    radiative, radial:      3/2 sum n(n+1)(2n+1) |(j_n(y) - a_n h_n(y)) / y|^2
    radiative, tangential:  3/4 sum (2n+1) [|j_n(y) - b_n h_n(y)|^2 + |(psi_n'(y) - a_n xi_n'(y)) / y|^2]

each relative to the dipole in the host alone; a dipole at angle t to the radial direction takes cos^2 t of the
radial and sin^2 t of the tangential rate. a_n is carried as (j_n(x) / h_n(x)) times a ratio of logarithmic
derivatives, and every product a_n h_n(y)^k is formed as j_n(x) h_n(y) (h_n(y) / h_n(x)), so the terms keep
their size ((r/d)^{2n}) where j_n(x) and h_n(y) alone are far apart in magnitude.

Pinned in tests/test_oracle_pins.py: a lossless sphere's total rate equals its radiative rate for both
orientations (the optical theorem; a sign slip between the two forms fails it); far from the sphere both tend to
1; a small lossy sphere quenches as 1 + c R^3 Im((eps - 1)/(eps + 2)) / (k^3 d^6) with c = 6 (radial) and 3/2
(tangential), the error falling as (R/d)^2.
"""

from __future__ import annotations

import numpy as np
from scipy.special import spherical_jn, spherical_yn


def _h(n, z, derivative=False):
    return spherical_jn(n, z, derivative) + 1j * spherical_yn(n, z, derivative)


def dipole_emission(radius: float, eps_sphere, wavelengths, n_host: float, d: float, cos2: float,
                    n_max: int | None = None) -> tuple[np.ndarray, np.ndarray]:
    """(total, radiative) decay-rate enhancements per wavelength of a dipole d from the centre of a sphere of
    permittivity eps_sphere (scalar or per wavelength) in a host of index n_host; cos2 = cos^2 of the dipole's
    angle to the radial direction."""
    wl = np.atleast_1d(np.asarray(wavelengths, dtype=float))
    eps = np.broadcast_to(np.asarray(eps_sphere, dtype=complex), wl.shape)
    if d <= radius:
        raise ValueError("the dipole must sit outside the sphere")
    tot, rad = [], []
    with np.errstate(all="ignore"):  # the spherical functions of x leave the float range past convergence
        for lam, e in zip(wl, eps):
            t_, r_ = _rates(radius, lam, e, n_host, d, cos2, n_max)
            tot.append(t_)
            rad.append(r_)
    return np.array(tot), np.array(rad)


def _rates(radius, lam, e, n_host, d, cos2, n_max):
    """(total, radiative) at one wavelength."""
    k = 2 * np.pi * n_host / lam
    x, y = k * radius, k * d
    m = np.sqrt(e) / n_host
    # terms fall as (r/d)^{2n}: 16 / ln(d/r) orders reach 1e-14, and the Mie terms need ~y + 4 y^(1/3)
    N = n_max or int(max(30, y + 4 * y ** (1 / 3) + 10, 16 / np.log(d / radius) + 10))
    n = np.arange(1, N + 1)
    jx, jpx = spherical_jn(n, x), spherical_jn(n, x, True)
    hx, hpx = _h(n, x), _h(n, x, True)
    jm, jpm = spherical_jn(n, m * x), spherical_jn(n, m * x, True)
    d_psi_x = 1 / x + jpx / jx  # psi_n'(x) / psi_n(x)
    d_xi_x = 1 / x + hpx / hx
    d_psi_m = 1 / (m * x) + jpm / jm
    A = (m * d_psi_x - d_psi_m) / (m * d_xi_x - d_psi_m)  # a_n = (j_n(x) / h_n(x)) A
    B = (d_psi_x - m * d_psi_m) / (d_xi_x - m * d_psi_m)
    jy, jpy = spherical_jn(n, y), spherical_jn(n, y, True)
    hy, hpy = _h(n, y), _h(n, y, True)
    xi_p = hy + y * hpy  # xi_n'(y)
    psi_p = jy + y * jpy
    scale_h = jx * hy * (hy / hx)  # j_n(x) h_n(y)^2 / h_n(x)
    a_hy = jx * (hy / hx) * A  # a_n h_n(y)
    b_hy = jx * (hy / hx) * B
    a_xi = jx * (xi_p / hx) * A  # a_n xi_n'(y)
    w_r = n * (n + 1) * (2 * n + 1)
    w_t = 2 * n + 1
    # the sphere's terms a_n h_n(y)^k: past the series' convergence the spherical functions of x leave the
    # float range (j_n(x) -> 0, h_n(x) -> inf) and the terms, far below 1e-14, are dropped
    sph_t = w_r * A * scale_h / y ** 2
    sph_p = w_t * (B * scale_h + a_xi * xi_p / y ** 2)
    ok = np.isfinite(sph_t) & np.isfinite(sph_p) & np.isfinite(a_hy) & np.isfinite(b_hy) & np.isfinite(a_xi)
    if not ok[: int(14 / np.log(d / radius))].all():
        raise FloatingPointError("spherical functions out of range before the series converged")
    sph_t, sph_p, a_hy, b_hy, a_xi = (np.where(ok, v, 0) for v in (sph_t, sph_p, a_hy, b_hy, a_xi))
    t_perp = 1 - 1.5 * np.real(np.sum(sph_t))
    t_par = 1 - 0.75 * np.real(np.sum(sph_p))
    r_perp = 1.5 * np.sum(w_r * np.abs((jy - a_hy) / y) ** 2)
    r_par = 0.75 * np.sum(w_t * (np.abs(jy - b_hy) ** 2 + np.abs((psi_p - a_xi) / y) ** 2))
    return cos2 * t_perp + (1 - cos2) * t_par, cos2 * r_perp + (1 - cos2) * r_par
