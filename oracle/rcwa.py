"""Rigorous coupled-wave analysis of lamellar (1D) gratings in planar incidence (TEST INFRASTRUCTURE ONLY: the
whole-solve check of the periodic driver's diffraction orders, as the reference checks its FDTD against its RCWA in
test_fdtd_periodic.py:141-192).

The reference's RCWA (backends/rcwa.py, backends/_rcwa/) is a general 2D-lattice S-matrix code; this restates the
published method for the case the checks need -- a stack of layers, each with boxes of constant permittivity along
x, period L, incidence in the xz plane (azimuth 0) -- after Moharam, Grann, Pommet and Gaylord, JOSA A 12, 1068
(1995), with Li's factorisation rule for TM (JOSA A 13, 1870 (1996)), written for the e^{-i omega t} convention:

* harmonics n = -N..N, kx_n / k0 = n_sup sin(theta) + n lambda / L; eps(x) = sum_m eps_m e^{i m 2 pi x / L}, box
  (centre c, width w) coefficients (w / L) sinc(m w / L) e^{-i 2 pi m c / L}; E = [eps_{m-n}], A = [(1/eps)_{m-n}];
* with z' = k0 z, a layer's tangential field f (TE: E_y; TM: eta_0 H_y) obeys f'' = -B f, B = E - Kx^2 (TE) or
  A^{-1} (I - Kx E^{-1} Kx) (TM); B = W diag(gamma^2) W^{-1}, Im gamma >= 0; f = W (e^{i gamma (z' - z_top)} c+ +
  e^{-i gamma (z' - z_bot)} c-), and the other tangential field (TE: H_x, TM: E_x, up to one common factor) is
  V (e^{..} c+ - e^{..} c-) with V = W Gamma (TE) or A W Gamma (TM); half-spaces have W = I, V = Gamma (TE) or
  Gamma / eps (TM);
* both tangential fields are continuous at every interface; the incident order is 1 in order 0 of the superstrate,
  r and t are the outgoing amplitudes; all interfaces are solved as one linear system;
* R_n = |r_n|^2 Re(V_sup,n) / V_sup,0 and T_n = |t_n|^2 Re(V_sub,n) / V_sup,0.

Pinned in tests/test_oracle_pins.py: a uniform layer is the characteristic-matrix TMM (oracle/tmm.py) for TE and TM
at normal and oblique incidence with every other order empty; lossless gratings conserve energy; the box Fourier
coefficients equal the FFT of the sampled profile; a mirror-symmetric grating at normal incidence puts equal power
in orders +n and -n; specular reflection is reciprocal (R_0(theta) = R_0(-theta)) for an asymmetric grating; a
subwavelength grating is its effective medium (TE <eps>, TM 1 / <1/eps>); results converge in N.
"""

from __future__ import annotations

import numpy as np


def box_coefficients(eps_bg, boxes, period: float, m, inverse: bool = False) -> np.ndarray:
    """Fourier coefficients (index m) of eps(x) (or 1/eps(x) when ``inverse``): background eps_bg with
    non-overlapping boxes (centre, width, eps) within one period."""
    val = (lambda e: 1.0 / complex(e)) if inverse else (lambda e: complex(e))
    m = np.asarray(m)
    out = np.where(m == 0, val(eps_bg), 0.0).astype(complex)
    for c, w, e in boxes:
        out = out + (val(e) - val(eps_bg)) * (w / period) * np.sinc(m * w / period) \
            * np.exp(-2j * np.pi * m * c / period)
    return out


def _branch(g):
    return np.where(g.imag < 0, -g, g)


def _layer_modes(eps_bg, boxes, period, kx, pol):
    """(W, V, gamma) of a grating layer."""
    M = len(kx)
    idx = np.arange(M)
    d = idx[:, None] - idx[None, :]
    E = box_coefficients(eps_bg, boxes, period, d)
    Kx = np.diag(kx)
    if pol == "s":
        B = E - Kx @ Kx
        lam, W = np.linalg.eig(B)
        g = _branch(np.sqrt(lam.astype(complex)))
        return W, W * g[None, :], g
    A = box_coefficients(eps_bg, boxes, period, d, inverse=True)
    B = np.linalg.solve(A, np.eye(M) - Kx @ np.linalg.solve(E, Kx))
    lam, W = np.linalg.eig(B)
    g = _branch(np.sqrt(lam.astype(complex)))
    return W, A @ (W * g[None, :]), g


def lamellar(layers, eps_sup: float, eps_sub: complex, period: float, wl: float, theta_rad: float, pol: str,
             n_harm: int = 25):
    """(orders n, R_n, T_n) of a stack of grating layers [(thickness nm, eps background, [(centre, width, eps)])]
    between a lossless superstrate (light comes from it) and a substrate; pol "s" (TE, E along y) or "p" (TM)."""
    n = np.arange(-n_harm, n_harm + 1)
    M = len(n)
    kx = np.sqrt(eps_sup) * np.sin(theta_rad) + n * wl / period
    k0 = 2 * np.pi / wl

    def half_space(eps):
        g = _branch(np.sqrt(complex(eps) - kx.astype(complex) ** 2))
        return g if pol == "s" else g / eps

    v_sup, v_sub = half_space(eps_sup), half_space(eps_sub)
    modes = []  # per layer (W, V, gamma, e^{i gamma k0 h})
    for h, e, bx in layers:
        W, V, g = _layer_modes(e, bx, period, kx, pol)
        modes.append((W, V, g, np.exp(1j * g * k0 * h)))
    L = len(layers)
    nu = M * (2 + 2 * L)
    G = np.zeros((2 * M * (L + 1), nu), dtype=complex)
    rhs = np.zeros(2 * M * (L + 1), dtype=complex)
    col_r, col_t = 0, M * (1 + 2 * L)
    delta = (n == 0).astype(complex)
    I = np.eye(M)
    for i in range(L + 1):
        rs, rv = slice(2 * M * i, 2 * M * i + M), slice(2 * M * i + M, 2 * M * (i + 1))
        # the region above interface i, at its bottom face
        if i == 0:
            G[rs, col_r:col_r + M] += I
            rhs[rs] -= delta
            G[rv, col_r:col_r + M] += -np.diag(v_sup)
            rhs[rv] -= v_sup * delta
        else:
            W, V, _, X = modes[i - 1]
            cp, cm = M * (1 + 2 * (i - 1)), M * (2 + 2 * (i - 1))
            G[rs, cp:cp + M] += W * X[None, :]
            G[rs, cm:cm + M] += W
            G[rv, cp:cp + M] += V * X[None, :]
            G[rv, cm:cm + M] += -V
        # minus the region below it, at its top face
        if i == L:
            G[rs, col_t:col_t + M] -= I
            G[rv, col_t:col_t + M] -= np.diag(v_sub)
        else:
            W, V, _, X = modes[i]
            cp, cm = M * (1 + 2 * i), M * (2 + 2 * i)
            G[rs, cp:cp + M] -= W
            G[rs, cm:cm + M] -= W * X[None, :]
            G[rv, cp:cp + M] -= V
            G[rv, cm:cm + M] -= -V * X[None, :]
    u = np.linalg.solve(G, rhs)
    r, t = u[col_r:col_r + M], u[col_t:col_t + M]
    v0 = float(np.real(v_sup[n == 0][0]))
    return n, np.abs(r) ** 2 * np.real(v_sup) / v0, np.abs(t) ** 2 * np.real(v_sub) / v0
