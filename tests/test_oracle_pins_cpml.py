"""Pins of the oracle's CPML with the complex-frequency-shifted terms switched on (kappa > 1, alpha > 0).

The CFS stretch 1/s = 1/kappa - (sigma/kappa^2) / (i w + sigma/kappa + alpha) (Roden & Gedney, MOTL 27, 334
(2000); Taflove & Hagness 3rd ed. sec. 7.9) turns the stretched derivative into

    (1/s) d = d + (1/kappa - 1) d + xi * d,    xi(t) = -(sigma/kappa^2) exp(-(sigma/kappa + alpha) t),

and the piecewise-constant recursive convolution samples xi * d as sum_m Z(m) d^{n-m}, Z(m) the integral of xi
over [m dt, (m+1) dt]. The first test drives the oracle's psi recursion (engine.py:636-663 as restated in
oracle/fdtd.py) with a known sequence of differences and compares the field increments with that convolution,
Z(m) found by numerical quadrature of xi (not from the closed-form b, c), sigma, kappa, alpha per node from the
published polynomial grading (sigma_max rho^m, 1 + (kappa_max - 1) rho^m, alpha_max (1 - rho)). A plausible slip --
the alpha grading reversed, the sign of 1/kappa - 1, a dropped kappa in c -- fails it (checked below by planting each
slip in the oracle). The second test is the brute-force reflection test at kappa_max = 2, alpha_max > 0.
"""

import numpy as np
import pytest
from scipy.integrate import quad

from oracle import fdtd, sources

DT = 0.95 / np.sqrt(3)


def _grading(spec, n_lo, n_hi, N, h, u):
    """Published CPML grading (Taflove & Hagness eqs. 7.60, 7.102, 7.79): rho the depth into the layer in
    units of its thickness, measured from its inner face (node n_lo, or N - 1 - n_hi) outwards."""
    u = np.asarray(u, dtype=float)
    out = []
    for x in u:
        if n_lo and x < n_lo:
            rho, eref = (n_lo - x) / n_lo, spec.eps_ref
        elif n_hi and x > N - 1 - n_hi:
            rho, eref = (x - (N - 1 - n_hi)) / n_hi, spec.eps_ref_high or spec.eps_ref
        else:
            rho, eref = 0.0, spec.eps_ref
        rho = min(rho, 1.0)
        sig = 0.8 * (spec.m + 1) / (np.sqrt(eref) * h) * rho ** spec.m
        kap = 1 + (spec.kappa_max - 1) * rho ** spec.m
        alp = spec.alpha_max * (1 - rho) if rho > 0 else 0.0
        out.append((sig, kap, alp))
    return out


def _Z(sig, kap, alp, dt, n):
    """Z(m), m < n: the CFS kernel xi integrated over each step, by adaptive quadrature."""
    rate = sig / kap + alp
    xi = lambda t: -(sig / kap ** 2) * np.exp(-rate * t)  # noqa: E731
    return np.array([quad(xi, m * dt, (m + 1) * dt, epsabs=0, epsrel=1e-13)[0] for m in range(n)])


def _cfs_mismatch(plant=None) -> float:
    """Largest relative mismatch between the oracle's E-side CPML increments along z and the brute-force
    CFS convolution, for a random sequence of H_y patterns (E_x += cb (-(1/s_z) dH_y/dz))."""
    shape, nz_pml, steps = (3, 3, 22), 6, 40
    spec = fdtd.PMLSpec(cells=((0, 0), (0, 0), (nz_pml, nz_pml)), kappa_max=2.5, alpha_max=0.35, m=3.0,
                        eps_ref=1.7)
    eng = fdtd.Engine(shape=shape, dx=1.0, dt=DT, inv_eps=np.ones((3,) + shape), pml=spec)
    if plant == "kappa_sign":
        eng.pml_E = [t[:6] + (None if t[6] is None else -t[6],) + t[7:] for t in eng.pml_E]
    rng = np.random.default_rng(7)
    seq = rng.standard_normal((steps, shape[2]))
    out = np.zeros((steps, shape[2]))
    for n in range(steps):
        eng.E[:] = 0
        eng.H[:] = 0
        eng.H[1][1, 1, :] = seq[n]
        eng.update_E()
        out[n] = eng.E[0][1, 1, :]
    # expected: E_x(k) = cb * -( (1/kappa) D^n + sum_m Z(m) D^{n-m} ),  D^n = H_y(k) - H_y(k-1)
    D = np.zeros_like(seq)
    D[:, 1:] = seq[:, 1:] - seq[:, :-1]
    cb = DT
    N = shape[2]
    worst = 0.0
    for k in range(1, N - 1):
        (sig, kap, alp), = _grading(spec, nz_pml, nz_pml, N, 1.0, [k])
        exp = -cb * D[:, k] / kap
        if sig > 0:
            Z = _Z(sig, kap, alp, DT, steps)
            exp = exp - cb * np.array([np.dot(Z[:n + 1], D[n::-1, k]) for n in range(steps)])
        worst = max(worst, np.abs(out[:, k] - exp).max() / np.abs(exp).max())
    return worst


def test_cfs_cpml_is_the_recursive_convolution_of_the_stretch():
    assert _cfs_mismatch() < 1e-12


def _slipped_profile(which):
    def profile(spec, n_lo, n_hi, N, h, dt, u):
        u = np.asarray(u, dtype=float)
        s_lo = 0.8 * (spec.m + 1) / (np.sqrt(spec.eps_ref) * h)
        s_hi = 0.8 * (spec.m + 1) / (np.sqrt(spec.eps_ref_high or spec.eps_ref) * h)
        r_lo = np.clip((n_lo - u) / n_lo, 0, 1) if n_lo else np.zeros_like(u)
        r_hi = np.clip((u - (N - 1 - n_hi)) / n_hi, 0, 1) if n_hi else np.zeros_like(u)
        rho = np.maximum(r_lo, r_hi)
        sig = np.where(r_lo >= r_hi, s_lo, s_hi) * rho ** spec.m
        kap = 1 + (spec.kappa_max - 1) * rho ** spec.m
        alp = spec.alpha_max * (rho if which == "alpha_rho" else 1 - rho) * (rho > 0)
        b = np.exp(-(sig / kap + alp) * dt)
        den = sig * kap + (kap if which == "c_kappa" else kap ** 2) * alp
        c = np.where(sig > 0, sig * (b - 1) / np.where(den > 0, den, 1), 0.0)
        return b, c, kap
    return profile


@pytest.mark.parametrize("which", ["alpha_rho", "c_kappa"])
def test_cfs_pin_fails_on_a_planted_grading_slip(monkeypatch, which):
    monkeypatch.setattr(fdtd, "cpml_profile", _slipped_profile(which))
    assert _cfs_mismatch() > 1e-3


def test_cfs_pin_fails_on_a_planted_kappa_sign():
    assert _cfs_mismatch(plant="kappa_sign") > 1e-3


def _dipole_run(shape, pml, pos, steps, probe, kappa_max, alpha_max):
    inv = np.ones((3,) + shape)
    src = sources.PointDipole(shape=shape, origin=np.zeros(3), dx=1.0, dt=DT, pos=pos, pdir=(0, 0, 1.0),
                              inv_eps=inv, omega_c=0.5, omega_half=0.25, max_steps=steps + 5)
    spec = fdtd.PMLSpec(cells=((pml, pml),) * 3, kappa_max=kappa_max, alpha_max=alpha_max) if pml else None
    eng = fdtd.Engine(shape=shape, dx=1.0, dt=DT, inv_eps=inv, pml=spec, sources=[src])
    out = []
    for _ in range(steps):
        eng.step()
        out.append(eng.E[2][probe])
    return np.asarray(out)


def test_cfs_cpml_reflection_is_small_against_a_large_domain():
    """The brute-force reflection test of the plain CPML, rerun with kappa_max = 2 and alpha_max = 0.05 w_c (the
    drivers' choice, 0.05 * 2 pi n / lambda_max): a probe 2 cells from an 8-cell CFS layer against the same probe
    in a domain whose walls have not echoed yet."""
    small, big = (32, 32, 32), (78, 78, 78)
    steps = 150
    a = _dipole_run(small, 8, (16.0, 16.0, 16.2), steps, (10, 16, 16), 2.0, 0.025)
    b = _dipole_run(big, 0, (39.0, 39.0, 39.2), steps, (33, 39, 39), 1.0, 0.0)
    rel = np.abs(a - b).max() / np.abs(b).max()
    assert rel < 5e-3
