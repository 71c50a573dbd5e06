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


def throughput_engine(scene: dict, device="cuda:0", max_steps: int = 10000, tiling=None, material: str = "auto",
                      fused: bool | None = None):
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
                     material=material, fused=fused)
    finally:
        Engine.tiling = old
    return eng


# --- batched oblique sweep (periodic.py:392-459, 503-519) --------------------------------------------

def oblique_sweep_scene(n_inst: int = 64, nz: int = 256, size_lo: int = 48, size_hi: int = 96, seed: int = 2,
                        pattern_seed: int = 3, d: float = 10.0, pml: int = 16, gap: int = 40, pad: int = 8,
                        layer: int = 40, eps_pillar: float = 12.1, eps_sub: float = 2.1, courant: float = 0.99,
                        n_angles: int = 8, n_wl: int = 8) -> list:
    """Independent runs of an oblique sweep over periodic metasurface cells: angles linspace(0, 85, n_angles)
    deg x wavelengths linspace(500, 1500, n_wl) nm, polarisation alternating s/p by run index. Each run has
    its own unit cell of nx, ny ~ U{size_lo..size_hi} real cells (seed) plus ghost planes, nz nodes along z
    with a `pml`-cell CPML at both ends, a binary pillar pattern (eps 1.0 / eps_pillar on 4 x 4-cell blocks,
    p = 0.5, pattern_seed) in a `layer`-cell layer on an eps_sub substrate, the reference's z layout
    (k_r = pml + gap, k_src = k_r + 2, k_top = k_src + pad, k_t = k_bot + 3; periodic.py:181-186), its time
    step, DFT stride and pulse half-width (periodic.py:304-331, 418-421). Angle 0 runs at normal incidence
    (real fields, a 1D-fed TFSF plane); the others at fixed k_par with a Bloch plane source (complex
    fields)."""
    rng = np.random.default_rng(seed)
    prng = np.random.default_rng(pattern_seed)
    angles = np.linspace(0.0, 85.0, n_angles)
    wls = np.linspace(500.0, 1500.0, n_wl)
    k_r = pml + gap
    k_src = k_r + 2
    k_top = k_src + pad
    k_bot = k_top + layer
    k_t = k_bot + 3
    if k_t + 2 >= nz - pml:
        raise ValueError("nz too small for the layout")
    runs = []
    for q in range(n_inst):
        theta, lam = float(angles[(q // n_wl) % n_angles]), float(wls[q % n_wl])
        pol = "s" if q % 2 == 0 else "p"
        nx, ny = (int(v) for v in rng.integers(size_lo, size_hi + 1, size=2))
        shape = (nx + 2, ny + 2, nz)
        blocks = prng.random(((nx + 3) // 4, (ny + 3) // 4)) < 0.5
        pattern = np.kron(blocks, np.ones((4, 4), bool))[:nx, :ny]
        eps = np.ones(shape)
        eps[1:nx + 1, 1:ny + 1, k_top:k_bot] = np.where(pattern[:, :, None], eps_pillar, 1.0)
        eps[:, :, k_bot:] = eps_sub
        inv_eps = np.broadcast_to(1.0 / eps, (3,) + shape).copy()
        omega = 2 * np.pi / lam
        k = omega * math.sin(math.radians(theta))  # superstrate eps = 1
        kpar = (k, 0.0)
        inv2 = 3.0 / d ** 2  # both periodic axes have more than one cell (periodic.py:304-314)
        dt = courant / math.sqrt(inv2)
        gaps = [abs(omega - k / math.sqrt(e)) for e in (1.0, eps_sub)]
        half = float(np.clip(min(gaps) / 4, 0.01 * omega, 0.15 * omega))  # periodic.py:324-331
        stride = max(1, int(2 * np.pi / (2 * omega + 4 * half) / dt))  # periodic.py:421
        runs.append({"shape": shape, "spacing": (d, d, d), "inv_eps": inv_eps, "theta": theta, "wavelength": lam,
                     "pol": pol, "kpar": kpar, "dt": dt, "omega": omega, "omega_half": half, "stride": stride,
                     "k_r": k_r, "k_src": k_src, "k_top": k_top, "k_t": k_t, "pml": pml, "eps_sub": eps_sub,
                     "n_cells": (nx, ny)})
    return runs


def common_stride(runs: list) -> int:
    """A DFT stride every run of a batch can use: the smallest run stride (each run's stride is its
    aliasing-safe maximum, so a smaller one is valid for it)."""
    return min(r["stride"] for r in runs)


def oblique_engine(run: dict, device="cuda:0", max_steps: int = 4000, stride: int | None = None):
    """The HIP engine of one sweep run with its source and the reflection, transmission and incident
    monitors (periodic.py:392-459)."""
    from .engine import DFTMonitor, Engine, PMLSpec
    from .run import plane_points
    from .sources import BlochPlaneSource, TFSFPlaneWave

    shape, d = run["shape"], run["spacing"]
    lam_max = run["wavelength"]
    pml = PMLSpec(cells=((0, 0), (0, 0), (run["pml"], run["pml"])), alpha_max=0.05 * 2 * np.pi / lam_max,
                  eps_ref=1.0, eps_ref_high=run["eps_sub"])
    stride = stride or run["stride"]
    om = np.array([run["omega"]])
    mon_r = DFTMonitor(om, *plane_points(shape, run["n_cells"], run["k_r"]), energy=1)
    mon_t = DFTMonitor(om, *plane_points(shape, run["n_cells"], run["k_t"]), energy=1)
    if run["theta"] == 0.0:
        phi = 0.0
        ehat = np.array([-np.sin(phi), np.cos(phi), 0.0]) if run["pol"] == "s" else np.array([1.0, 0.0, 0.0])
        origin = np.array([-(shape[0] - 2) * d[0] / 2 - d[0], -(shape[1] - 2) * d[1] / 2 - d[1], -run["k_top"] * d[2]])
        r0 = origin + np.array([1, 1, run["k_top"]]) * np.asarray(d)
        big = 1 << 30
        src = TFSFPlaneWave(shape=shape, origin=origin, dx=d, dt=run["dt"], r0=r0, lo=np.array([-big, -big, run["k_src"]]),
                            hi=np.array([big, big, big]), khat=np.array([0.0, 0.0, 1.0]), ehat=ehat, eps_b=1.0,
                            omega_c=run["omega"], omega_half=run["omega_half"], max_steps=max_steps,
                            periodic=(True, True, False))
        mon_i = DFTMonitor(om, *src.probe(), source=src)
        bloch = (0.0, 0.0, 0.0)
    else:
        col_pml = PMLSpec(cells=pml.cells, alpha_max=pml.alpha_max, eps_ref=1.0)
        src = BlochPlaneSource(shape=shape, spacing=d, dt=run["dt"], k_src=run["k_src"], k_launch=run["k_src"] - 1,
                               kpar=run["kpar"], pol=run["pol"], eps_b=1.0, pml=col_pml, omega_c=run["omega"],
                               omega_half=run["omega_half"], max_steps=max_steps)
        mon_i = DFTMonitor(om, *src.probe(run["k_top"]), source=src)
        bloch = (run["kpar"][0], run["kpar"][1], 0.0)
    return Engine(device, shape=shape, dx=d, dt=run["dt"], single=True, inv_eps=run["inv_eps"], groups=[], pml=pml,
                  sources=[src], monitors=[mon_r, mon_t, mon_i], dft_stride=stride, periodic=(True, True, False),
                  bloch=bloch)


# --- adjoint inverse design (an SOI-like waveguide with a design region) ---------------------------------

def adjoint_scene(shape=(256, 192, 96), dx: float = 20.0, design=(80, 80, 11), seed: int = 4, pml: int = 16,
                  n_steps: int = 3000, stride: int = 25, wavelength: float = 1550.0, courant: float = 0.99) -> dict:
    """A silicon strip waveguide (eps 3.48^2, 500 x 220 nm: 25 x 11 cells) along x in oxide cladding
    (eps 1.44^2), broken in the middle by a design region of `design` cells (the waveguide's thickness) whose
    density rho ~ U[0, 1] (seed) sets eps = 1.44^2 + rho (3.48^2 - 1.44^2); 16-cell CPML on every face; a soft
    mode-injection source of E_y on an x plane near the input (a Gaussian approximation of the fundamental TE
    profile over the core, Gaussian pulse at `wavelength` with half-width 0.1 omega); a DFT monitor of E_y, E_z
    and face-averaged H on an x plane near the output at `wavelength`. The objective is the power in the same
    profile at the output, |sum_p w_p E_y(p)|^2. DFT stride 25 (finer than the reference rule's 90 at this dt;
    it divides the 3000 steps and the 100-step checkpoint interval)."""
    n0, n1, n2 = shape
    e_clad, e_core = 1.44 ** 2, 3.48 ** 2
    eps = np.full(shape, e_clad)
    jc, kc = n1 // 2, n2 // 2
    wy, wz = 25, 11
    j_lo, k_lo = jc - wy // 2, kc - wz // 2
    eps[:, j_lo:j_lo + wy, k_lo:k_lo + wz] = e_core
    di, dj, dk = design
    lo = (n0 // 2 - di // 2, jc - dj // 2, k_lo)
    hi = (lo[0] + di, lo[1] + dj, lo[2] + dk)
    rho = np.random.default_rng(seed).uniform(0.0, 1.0, size=design)
    inv_eps = np.broadcast_to(1.0 / eps, (3,) + shape).copy()
    dt = courant * dx / math.sqrt(3)
    omega = 2 * np.pi / wavelength
    # E_y window around the core: +-1 core width in y, +-2 thicknesses in z
    jj, kk = np.meshgrid(np.arange(j_lo - wy, j_lo + 2 * wy), np.arange(k_lo - 2 * wz, k_lo + 3 * wz), indexing="ij")
    jj, kk = jj.ravel(), kk.ravel()
    prof = np.exp(-((jj - (j_lo + wy / 2 - 0.5)) / (0.45 * wy)) ** 2 - ((kk - (k_lo + wz / 2 - 0.5)) / (0.6 * wz)) ** 2)
    i_src, i_mon = pml + 14, n0 - pml - 14
    e_src = flat_index(1, np.full(jj.size, i_src), jj, kk, shape)
    tau = 2.0 / (0.1 * omega)
    t = (np.arange(n_steps + 2) + 0.5) * dt
    pulse = np.exp(-((t - 3.5 * tau) / tau) ** 2) * np.sin(omega * (t - 3.5 * tau))
    im = np.full(jj.size, i_mon)
    e_mon = np.concatenate([flat_index(1, im, jj, kk, shape), flat_index(2, im, jj, kk, shape)])
    h1 = np.concatenate([flat_index(2, im - 1, jj, kk, shape), flat_index(1, im - 1, jj, kk, shape)])
    h2 = np.concatenate([flat_index(2, im, jj, kk, shape), flat_index(1, im, jj, kk, shape)])
    w = np.concatenate([prof, np.zeros_like(prof)]) / np.linalg.norm(prof)
    return {"shape": shape, "dx": dx, "dt": dt, "inv_eps": inv_eps, "pml": pml, "lo": lo, "hi": hi, "eps_lo": e_clad,
            "eps_hi": e_core, "rho": rho, "n_steps": n_steps, "stride": stride, "omega": np.array([omega]),
            "e_src": e_src, "m_src": prof[:, None], "t_src": pulse[:, None], "e_mon": e_mon, "h1": h1, "h2": h2,
            "w": w, "n_cells": int(np.prod(shape))}


def adjoint_problem(scene: dict, device="cuda:0", **kw):
    """The AdjointProblem of an adjoint scene (CPML alpha from the source frequency)."""
    from .adjoint import AdjointProblem, DesignRegion
    from .engine import DFTMonitor, PMLSpec
    from .sources import ModeInjection

    p = scene["pml"]
    src = ModeInjection(np.zeros(0, np.int64), np.zeros((0, 1)), np.zeros((scene["n_steps"] + 2, 1)), scene["e_src"],
                        scene["m_src"], scene["t_src"])
    mon = DFTMonitor(scene["omega"], scene["e_mon"], scene["h1"], scene["h2"])
    pml = PMLSpec(cells=((p, p),) * 3, alpha_max=0.05 * float(scene["omega"][0]), eps_ref=scene["eps_lo"])
    return AdjointProblem(device, shape=scene["shape"], dx=scene["dx"], dt=scene["dt"], inv_eps=scene["inv_eps"],
                          pml=pml, sources=[src], monitor=mon,
                          design=DesignRegion(scene["lo"], scene["hi"], scene["eps_lo"], scene["eps_hi"]),
                          n_steps=scene["n_steps"], dft_stride=scene["stride"], groups=scene.get("groups", ()), **kw)


def mode_power(acc_e, w):
    """|sum_p w_p acc_e[0, p]|^2 (the adjoint scene's objective; numpy or torch)."""
    if isinstance(acc_e, np.ndarray):
        return abs(np.sum(w * acc_e[0])) ** 2
    import torch
    return (torch.as_tensor(w, device=acc_e.device).to(acc_e.dtype) * acc_e[0]).sum().abs() ** 2


# --- plain Au sphere (the whole-solve workload) ---------------------------------------------------------------------

# Au as eps_inf + Drude + 2 Lorentz poles over 400-800 nm (energies in eV): a fixed parameter set of this workload,
# close to measured gold (eps(600 nm) ~ -11.6 + 1.6i, interband onset near 2.4 eV). The poles follow from these
# parameters through the reference's Drude-Lorentz -> pole map (media.py:68-75).
AU_DRUDE_2LORENTZ = {"eps_inf": 1.53, "wp_ev": 8.55, "gamma_ev": 0.0735,
                     "poles": [{"delta_eps": 0.94, "omega0_ev": 2.88, "gamma_ev": 0.74},
                               {"delta_eps": 1.36, "omega0_ev": 3.90, "gamma_ev": 1.25}]}


def au_sphere_sim(num_wavelengths: int = 81, run_time_fs: float | None = None, **options):
    """The plain Au sphere: r = 40 nm in water (n = 1.33), Au as Drude + 2 Lorentz poles, plane wave along +z
    polarised along x, 81 wavelengths 400-800 nm, 2 nm cells, a 12-cell CPML, courant 0.99, mirror symmetry off.
    The reference lays this out as 97^3 nodes (96^3 cells; setup.py:640 always gives an odd node count). A
    `lightsmith_amd.problem.Simulation`, so it also runs through ``api.solve``."""
    from ..problem import DrudeLorentz, Particle, Simulation, Sphere, WavelengthRange
    au = DrudeLorentz(**AU_DRUDE_2LORENTZ)
    opts = {"spacing_nm": 2.0, "pml_cells": 12, "padding_nm": 12.0, "pml_gap_nm": 12.0, "courant": 0.99,
            "symmetry": "off", **({"run_time_fs": run_time_fs} if run_time_fs else {}), **options}
    return Simulation(particles=[Particle(shape=Sphere(radius_nm=40.0), material=au)], medium_index=1.33,
                      wavelengths=WavelengthRange(start_nm=400.0, stop_nm=800.0, num=num_wavelengths),
                      options={"fdtd": opts})


def au_sphere_scene(device="cuda:0", **kw):
    """The scattering Scene of au_sphere_sim (materials voxelised on `device`) and its options."""
    from ..problem import FDTDOptions, options_for
    from . import scattering
    sim = au_sphere_sim(**kw)
    opts = FDTDOptions.model_validate(options_for(sim, "lightsmith-fdtd", "fdtd"))
    return sim, opts, scattering.setup(sim, opts, device)


# --- batched adjoint with dispersion (multi-instance inverse design) -------------------------------------------------

# Ag as Drude + 3 Lorentz poles (energies in eV): a fixed parameter set of this workload.
AG_DRUDE_3LORENTZ = {"eps_inf": 1.0, "wp_ev": 9.01, "gamma_ev": 0.048,
                     "lorentz": [(0.79, 4.43, 0.65), (0.35, 5.10, 1.40), (0.60, 8.00, 2.50)]}


def batched_adjoint_scenes(n_inst: int = 16, seed: int = 6, n_steps: int = 2500, dx: float = 20.0, pml: int = 12,
                           stride: int = 25) -> list:
    """The multi-instance inverse-design workload: n_inst adjoint scenes of seeded uneven sizes
    (128-256 x 96-192 x 64-96 cells, seed 6), each the waveguide scene of adjoint_scene (a strip waveguide along x
    broken by a seeded density design region of up to 60 x 60 x 11 cells) with an Ag nanostructure -- a 6-cell
    cube of Drude + 3 Lorentz silver (pure D-form nodes) on the waveguide axis between the design region and the
    output monitor; dt from the reference's rule with the dispersive stability limit (setup.py:78-91); 12-cell
    CPML; n_steps steps, DFT every 25 steps."""
    rng = np.random.default_rng(seed)
    from .engine import DispersiveGroup
    from .media import drude_lorentz
    ag = drude_lorentz(AG_DRUDE_3LORENTZ["eps_inf"], AG_DRUDE_3LORENTZ["wp_ev"], AG_DRUDE_3LORENTZ["gamma_ev"],
                       AG_DRUDE_3LORENTZ["lorentz"], "Ag")
    dt = time_step(dx, 0.99, [Medium(1.44 ** 2), Medium(3.48 ** 2), ag])
    out = []
    for q in range(n_inst):
        shape = (int(rng.integers(128, 257)), int(rng.integers(96, 193)), int(rng.integers(64, 97)))
        design = (min(60, shape[0] // 4), min(60, shape[1] // 3), 11)
        sc = adjoint_scene(shape=shape, dx=dx, design=design, seed=int(rng.integers(1 << 30)), pml=pml,
                           n_steps=n_steps, stride=stride)
        sc["dt"] = dt
        sc["t_src"] = _mode_pulse(n_steps, dt, float(sc["omega"][0]))
        n0, n1, n2 = shape
        # the Ag cube: 6 cells on the waveguide axis, halfway between the design region and the monitor plane
        i_mon = n0 - pml - 14
        ic = (sc["hi"][0] + i_mon) // 2
        jc, kc = n1 // 2, n2 // 2
        box = [np.arange(ic - 3, ic + 3), np.arange(jc - 3, jc + 3), np.arange(kc - 3, kc + 3)]
        ii, jj, kk = (a.ravel() for a in np.meshgrid(*box, indexing="ij"))
        idx = np.sort(np.concatenate([flat_index(c, ii, jj, kk, shape) for c in range(3)]))
        sc["inv_eps"].reshape(-1)[idx] = 1.0
        sc["groups"] = [DispersiveGroup(idx=idx, media=[ag], frac=np.ones((1, len(idx))), w=None)]
        sc["ag_cells"] = int(len(ii))
        out.append(sc)
    return out


def _mode_pulse(n_steps: int, dt: float, omega: float) -> np.ndarray:
    """The mode source's Gaussian pulse (half-width 0.1 omega) sampled at the source's half steps."""
    tau = 2.0 / (0.1 * omega)
    t = (np.arange(n_steps + 2) + 0.5) * dt
    return (np.exp(-((t - 3.5 * tau) / tau) ** 2) * np.sin(omega * (t - 3.5 * tau)))[:, None]
