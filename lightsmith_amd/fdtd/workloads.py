"""Synthetic FDTD workloads with seeded inputs (NumPy, float64 host data).

``throughput_scene``: the plain-FDTD throughput configuration -- an n^3 grid (320^3 by default) with a
16-cell CPML on every face, permittivity from 200 seeded random spheres (radius 4-20 cells, eps uniform in
[1, 12]; staircase: each E_c position takes the first sphere that contains it, as the reference's
particle_slot_fn lets the first particle win overlaps), 10 % of the non-CPML cells dispersive with one
3-pole Lorentz ADE medium (eps_inf = 1; seeded omega_p, omega_0 in [1, 10] eV, gamma in [0.05, 0.5] eV), a
z-directed point dipole at the centre (Gaussian pulse at 1550 nm, half-width 0.15 omega_c), and one DFT
plane at z = n/2 + 32 (E_x, E_y and face-averaged H_x, H_y over the non-CPML (n - 32)^2 points, 16
wavelengths 1300-1800 nm). The spacing (20 nm) and courant number (0.95) are this workload's choices; dt
comes from the reference's rule with the dispersive stability limit (setup.py:78-91).
"""

from __future__ import annotations

import math

import numpy as np

from .media import HBARC_EV_NM, Medium, Pole, time_step


def flat_index(c, i, j, k, shape):
    n0, n1, n2 = shape
    return ((np.asarray(c) * n0 + np.asarray(i)) * n1 + np.asarray(j)) * n2 + np.asarray(k)


def lorentz_medium(seed: int, n_poles: int = 3) -> Medium:
    rng = np.random.default_rng(seed)
    poles = []
    for _ in range(n_poles):
        wp, w0, g = rng.uniform(1, 10), rng.uniform(1, 10), rng.uniform(0.05, 0.5)
        poles.append(Pole((wp / HBARC_EV_NM) ** 2, w0 / HBARC_EV_NM, g / HBARC_EV_NM))
    return Medium(1.0, tuple(poles), "lorentz3")


def throughput_scene(n: int = 320, pml: int = 16, n_spheres: int = 200, seed: int = 0, disp_frac: float = 0.1,
                     disp_seed: int = 1, pole_seed: int = 2, dx: float = 20.0, courant: float = 0.95,
                     n_freq: int = 16) -> dict:
    shape = (n, n, n)
    rng = np.random.default_rng(seed)
    eps = np.ones((3,) + shape)
    taken = np.zeros((3,) + shape, bool)
    spheres = []
    scale = n / 320.0
    for _ in range(n_spheres):
        r = max(1, int(round(rng.integers(4, 21) * scale)))
        lo, hi = pml + r, n - 1 - pml - r
        ctr = rng.uniform(lo, hi, size=3) if hi > lo else np.full(3, n / 2)
        e = rng.uniform(1, 12)
        spheres.append((ctr, r, e))
    for ctr, r, e in spheres:
        for c in range(3):
            off = np.zeros(3)
            off[c] = 0.5
            lo = np.maximum(np.floor(ctr - r - 1).astype(int), 0)
            hi = np.minimum(np.ceil(ctr + r + 1).astype(int), n)
            ax = [np.arange(lo[a], hi[a]) + off[a] - ctr[a] for a in range(3)]
            d2 = ax[0][:, None, None] ** 2 + ax[1][None, :, None] ** 2 + ax[2][None, None, :] ** 2
            inside = d2 <= r * r
            sub = (c, slice(lo[0], hi[0]), slice(lo[1], hi[1]), slice(lo[2], hi[2]))
            new = inside & ~taken[sub]
            eps[sub][new] = e
            taken[sub] |= inside
    inv_eps = 1.0 / eps
    drng = np.random.default_rng(disp_seed)
    disp = np.zeros(shape, bool)
    inner = (slice(pml, n - pml),) * 3
    disp[inner] = drng.random((n - 2 * pml,) * 3) < disp_frac
    cells = np.flatnonzero(disp.reshape(-1))
    inv_eps.reshape(3, -1)[:, cells] = 1.0  # D-form nodes (setup.py:187, 234)
    ii, jj, kk = np.unravel_index(cells, shape)
    disp_idx = np.concatenate([flat_index(c, ii, jj, kk, shape) for c in range(3)])
    med = lorentz_medium(pole_seed)
    media = [Medium(float(e), ()) for _, _, e in spheres] + [med]
    dt = time_step(dx, courant, media)
    omega_c = 2 * np.pi / 1550.0
    omega_half = 0.15 * omega_c
    wl = np.linspace(1300.0, 1800.0, n_freq)
    omega = 2 * np.pi / wl
    stride = max(1, int(2 * np.pi / (omega_c + 4 * omega_half + omega.max()) / dt))  # scattering.py:261
    # one DFT plane: E_x, E_y on node plane kz, H averaged over kz - 1/2 and kz + 1/2 (periodic.py:335-344)
    kz = n // 2 + 32 if n // 2 + 32 < n - pml else n // 2
    gi, gj = np.meshgrid(np.arange(pml, n - pml), np.arange(pml, n - pml), indexing="ij")
    gi, gj = gi.ravel(), gj.ravel()
    e_idx = np.concatenate([flat_index(0, gi, gj, kz, shape), flat_index(1, gi, gj, kz, shape)])
    h1 = np.concatenate([flat_index(1, gi, gj, kz - 1, shape), flat_index(0, gi, gj, kz - 1, shape)])
    h2 = np.concatenate([flat_index(1, gi, gj, kz, shape), flat_index(0, gi, gj, kz, shape)])
    return {"shape": shape, "dx": dx, "dt": dt, "inv_eps": inv_eps, "disp_idx": disp_idx, "medium": med,
            "pml": pml, "omega_c": omega_c, "omega_half": omega_half, "omega": omega, "stride": stride,
            "dipole_pos": np.full(3, (n // 2) * dx), "dipole_dir": (0.0, 0.0, 1.0), "monitor": (e_idx, h1, h2),
            "n_cells": int(np.prod(shape)), "n_dispersive_cells": int(len(cells)), "courant": courant}


def steps_for(total: int, stride: int) -> int:
    return int(math.ceil(total / stride)) * stride


def throughput_engine(scene: dict, device="cuda:0", max_steps: int = 10000, tiling=None, material: str = "auto"):
    """The HIP engine of a throughput scene: pure D-form group, CPML (alpha from the lowest frequency, as the
    emission driver sets it, emission.py:176-177), the dipole and the DFT plane (energy proxy 1)."""
    from .engine import DFTMonitor, DispersiveGroup, Engine, PMLSpec
    from .sources import PointDipole

    shape, idx = scene["shape"], scene["disp_idx"]
    groups = [DispersiveGroup(idx=idx, media=[scene["medium"]], frac=np.ones((1, len(idx))), w=None)] if len(idx) else []
    src = PointDipole(shape=shape, origin=np.zeros(3), dx=scene["dx"], dt=scene["dt"], pos=scene["dipole_pos"],
                      pdir=scene["dipole_dir"], inv_eps=scene["inv_eps"], omega_c=scene["omega_c"],
                      omega_half=scene["omega_half"], max_steps=max_steps)
    mon = DFTMonitor(scene["omega"], *scene["monitor"], energy=1)
    p = scene["pml"]
    pml = PMLSpec(cells=((p, p),) * 3, alpha_max=0.05 * float(np.min(scene["omega"])), eps_ref=1.0)
    old = Engine.tiling
    if tiling is not None:
        Engine.tiling = tuple(tiling)
    try:
        eng = Engine(device, shape=shape, dx=scene["dx"], dt=scene["dt"], single=True, inv_eps=scene["inv_eps"],
                     groups=groups, pml=pml, sources=[src], monitors=[mon], dft_stride=scene["stride"],
                     material=material)
    finally:
        Engine.tiling = old
    return eng
