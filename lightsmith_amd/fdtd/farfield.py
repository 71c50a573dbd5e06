"""Near-to-far transform over a closed flux box (setup.py:472-548 of the reference): the far-zone field radiated by
the box's equivalent surface currents J = n x H and M = -n x E (Love's equivalence), from the running DFTs of the
same samples the box's flux uses.

The surface is read off the flux box itself (``run.flux_points``): each sample pairs a tangential E component with
the face-averaged H of the other tangential component, and its flux coefficient is area x n . (e_E x e_H), so the
area is the coefficient's magnitude and the outward normal n = sign(coefficient) e_E x e_H; the sample sits at its E
node. With exp(-i omega t) fields, in units where the flux is 1/2 Re(E x H*) (so a medium of index n has wave
impedance 1/n), the field at distance r in direction u is e^{ikr}/r times

    E_far(u) = (-i k / 4 pi) [eta u x (u x N) + u x L],  N = sum J e^{-i k u.r'} dA,  L = sum M e^{-i k u.r'} dA,

and the power per solid angle n |E_far|^2 / 2.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np


@dataclass
class Surface:
    """Equivalent-current samples of a flux box: positions (nm), unit J and M directions per unit H and E sample
    amplitude, area weights (nm^2)."""

    pos: np.ndarray
    jdir: np.ndarray
    mdir: np.ndarray
    area: np.ndarray


def box_surface(e_idx, h1_idx, coeff, shape, origin, dx: float) -> Surface:
    """The Surface of a flux box given as run.flux_points returns it (reference flat indices)."""
    N = int(np.prod(shape))
    c_e, rest = np.divmod(np.asarray(e_idx, dtype=np.int64), N)
    c_h = np.asarray(h1_idx, dtype=np.int64) // N
    ijk = np.stack(np.unravel_index(rest, tuple(shape)), axis=1).astype(float)
    eye = np.eye(3)
    pos = np.asarray(origin, dtype=float) + (ijk + 0.5 * eye[c_e]) * dx
    coeff = np.asarray(coeff, dtype=float)
    nrm = np.sign(coeff)[:, None] * np.cross(eye[c_e], eye[c_h])
    return Surface(pos=pos, jdir=np.cross(nrm, eye[c_h]), mdir=-np.cross(nrm, eye[c_e]), area=np.abs(coeff))


def far_field(surf: Surface, e: np.ndarray, h: np.ndarray, directions: np.ndarray, k: float, eta: float,
              center=(0.0, 0.0, 0.0), chunk: int = 512) -> np.ndarray:
    """E_far (m, 3) in the unit directions (m, 3) at one frequency from the surface samples' DFTs e, h (n,), with
    the phase referred to ``center`` (nm)."""
    J = (h * surf.area)[:, None] * surf.jdir
    M = (e * surf.area)[:, None] * surf.mdir
    r = surf.pos - np.asarray(center, dtype=float)
    u_all = np.asarray(directions, dtype=float)
    out = np.empty((len(u_all), 3), dtype=complex)
    for a in range(0, len(u_all), chunk):
        u = u_all[a:a + chunk]
        ph = np.exp(-1j * k * (u @ r.T))
        Nv, Lv = ph @ J, ph @ M
        out[a:a + chunk] = (-1j * k / (4 * np.pi)) * (eta * np.cross(u, np.cross(u, Nv)) + np.cross(u, Lv))
    return out


def directions(theta_deg, phi_deg) -> np.ndarray:
    """Unit vectors (n_theta, n_phi, 3), theta from +z, phi from +x."""
    th = np.radians(np.asarray(theta_deg, dtype=float))[:, None]
    ph = np.radians(np.asarray(phi_deg, dtype=float))[None, :]
    return np.stack(np.broadcast_arrays(np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)), axis=-1)


__all__ = ["Surface", "box_surface", "directions", "far_field"]
