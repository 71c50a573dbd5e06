"""Plane-wave scattering by particles, solved end to end on the HIP engine (scattering.py:232-500 of the
reference, for a homogeneous host on a uniform grid):

    plan   -> cell size, grid layout, dt, DFT stride, step cap            (scattering.py:232-280)
    set-up -> ADE media, voxelised materials (subpixel / conformal),       (scattering.py:301-346)
              CPML, the table-fed plane wave, scattered- and total-field flux boxes
    run    -> run_until_decayed on the engine (HIP graphs per DFT chunk)  (setup.py:722-773)
    post   -> C_sca = outward flux of the scattered field / I, C_abs = -(net total-field flux - the background's
              own) / I, C_ext = C_sca + C_abs, I = n |E_inc|^2 / 2          (scattering.py:428-464)

A non-dispersive, lossless substrate below z = interface_z_nm is meshed apart as a planar background
(``voxel.planar_background``, the rules the source tables assume), the particles rasterised over it, and the plane
wave fed from tables of the grid's exact layered background (incident, reflected and transmitted), from above or
through the substrate, at any angle that does not graze the interface. The cross sections keep the same bookkeeping:
the background's own flux through the total-field box comes from the same tables.

The far field (dC_sca/dOmega) is the scattered-field box's near-to-far transform (``farfield``); the near field the
total E on a plane, interpolated from a slab of DFT nodes with the tables' background added outside the box; the
resonances the harmonic inversion (``resonance``) of probe fields sampled once per DFT chunk after the source is off.

Not on this path (the backend's ``supports`` declines them): dispersive or absorbing substrates, Gaussian beams,
the far field over a substrate, mirror symmetry (runs are full grids), graded grids,
tabulated materials (no Drude-Lorentz fit).
"""

from __future__ import annotations

import math
import time
from dataclasses import dataclass

import numpy as np
import torch

from . import farfield
from .engine import DFTMonitor, Engine, PMLSpec, TimeMonitor, filling_tiling, flat_index
from .layered import OMEGA_H, ObliqueLayeredPlaneWave
from .media import HBARC_EV_NM, Medium, Pole, time_step
from .run import FS, flux, flux_points, run_until_decayed
from .voxel import OnSubstrate, ParticleSlots, planar_background, voxelise


# --- media (media.py:60-129) --------------------------------------------------------------------------------------

def _from_drude_lorentz(eps_inf, wp_ev, gamma_ev, lorentz, label) -> Medium:
    poles = []
    if wp_ev > 0:
        poles.append(Pole((wp_ev / HBARC_EV_NM) ** 2, 0.0, gamma_ev / HBARC_EV_NM))
    for de, w0_ev, g_ev in lorentz:
        w0 = w0_ev / HBARC_EV_NM
        poles.append(Pole(de * w0 ** 2, w0, g_ev / HBARC_EV_NM))
    return Medium(float(eps_inf), tuple(poles), label)


def _matched_constant(eps: complex, omega_c: float, label: str) -> Medium:
    """A constant complex eps as a Drude term exact at omega_c (media.py:77-91): lossless eps >= 1 is a plain
    dielectric; otherwise eps_inf = max(1, eps' + eps''/20) and a Drude pole makes up the rest."""
    er, ei = eps.real, eps.imag
    if ei == 0 and er >= 1:
        return Medium(er, (), label)
    eps_inf = max(1.0, er + ei / 20)
    delta = eps_inf - er
    gamma = omega_c * ei / delta
    wp2 = omega_c ** 2 * (delta ** 2 + ei ** 2) / delta
    return Medium(eps_inf, (Pole(wp2, 0.0, gamma),), label)


def to_medium(material, wavelengths_nm, warnings: list) -> Medium:
    """An ADE medium for one material over the band (media.py:94-129)."""
    wl = np.asarray(wavelengths_nm, dtype=float)
    kind = getattr(material, "kind", type(material).__name__)
    if kind == "drude":
        return _from_drude_lorentz(material.eps_inf, material.wp_ev, material.gamma_ev, (), kind)
    if kind == "drude_lorentz":
        m = _from_drude_lorentz(material.eps_inf, material.wp_ev, material.gamma_ev,
                                [(p.delta_eps, p.omega0_ev, p.gamma_ev) for p in material.poles], kind)
        if any(p.sigma < 0 for p in m.poles):
            warnings.append("DrudeLorentz with negative oscillator strength: the FDTD update may be unstable")
        return m
    if kind in ("constant_index", "constant_permittivity"):
        eps = complex(material.eps(np.array([wl[0]]))[0])
        omega = 2 * np.pi / wl
        omega_c = 0.5 * (omega.min() + omega.max())
        m = _matched_constant(eps, omega_c, kind)
        if m.dispersive and wl.size > 1:
            dev = float(np.max(np.abs(m.eps(omega) - eps)) / abs(eps))
            if dev > 1e-3:
                warnings.append(f"{kind} eps={eps:.4g} is not causal as a constant; FDTD models it as a Drude term "
                                f"exact at {2 * np.pi / omega_c:.0f} nm (max deviation {dev:.1%} over the band). "
                                "Use a Drude/DrudeLorentz material for broadband accuracy.")
        return m
    raise ValueError(f"FDTD on this path does not support material kind {kind!r} (give Drude or DrudeLorentz "
                     "parameters; tabulated materials need a Drude-Lorentz fit)")


def particle_media(particles, medium_index: float, wl, warnings) -> tuple[list[Medium], list[list[int]]]:
    """ADE media (0 = the host) and each particle's region -> medium slots; equal materials share a medium
    (scattering.py:102-116)."""
    media = [Medium(medium_index ** 2, (), "medium")]
    keys: dict[str, int] = {}
    slot_of = []
    for p in particles:
        slots = []
        for m in p.region_materials():
            key = m.model_dump_json()
            if key not in keys:
                keys[key] = len(media)
                media.append(to_medium(m, wl, warnings))
            slots.append(keys[key])
        slot_of.append(slots)
    return media, slot_of


def band_wavelengths(sim) -> np.ndarray:
    """The wavelengths the pulse must carry: the results', the near field's and the far field's
    (scattering.py:135-140)."""
    wl = sim.wavelength_array()
    if sim.outputs.near_field is not None:
        wl = np.concatenate([wl, np.asarray(sim.outputs.near_field.wavelengths_nm, dtype=float)])
    if sim.outputs.far_field is not None:
        wl = np.concatenate([wl, np.asarray(sim.outputs.far_field.wavelengths_nm, dtype=float)])
    return wl


def scene_media(sim, warnings) -> tuple[list[Medium], list[list[int]]]:
    """The particles' media (``particle_media``) and, with a substrate, its medium as the last slot, fitted over
    the band's wavelengths (scattering.py:26-38)."""
    wl = band_wavelengths(sim)
    media, slot_of = particle_media(sim.particles, sim.medium_index, wl, warnings)
    if sim.substrate is not None:
        media.append(to_medium(sim.substrate.material, wl, warnings))
    return media, slot_of


def substrate_unsupported(sim) -> str:
    """Why this substrate problem is not on the HIP path ('' if it is; scattering.py:41-61)."""
    if sim.substrate is None:
        return ""
    k, _ = sim.illumination.unit_vectors()
    if abs(k[2]) < 0.1:
        return "with a substrate, the plane wave must not graze the interface (|direction_z| >= 0.1)"
    wl = sim.wavelength_array()
    eps = np.asarray(sim.substrate.material.eps(wl), dtype=complex)
    if np.any(np.abs(eps.imag) > 1e-4 * np.abs(eps)):
        return ("an absorbing substrate is not on the HIP path (its cross sections need the particles' own "
                "absorption and the far field)")
    if getattr(sim.substrate.material, "kind", "") not in ("constant_index", "constant_permittivity") \
            or np.any(eps.real < 1):
        return "the HIP path takes a non-dispersive substrate (constant index, eps >= 1)"
    return ""


# --- planning ------------------------------------------------------------------------------------------------------

def band(sim) -> tuple[float, float, float]:
    """(omega_c, omega_half, omega_max) of the pulse covering the wavelengths, the far field's included
    (scattering.py:143-152)."""
    om = 2 * np.pi / band_wavelengths(sim)
    omega_c = 0.5 * (om.min() + om.max())
    return float(omega_c), float(max(0.5 * (om.max() - om.min()), 0.15 * omega_c)), float(om.max())


def _n_max(materials, wl) -> tuple[float, bool]:
    n_max, metal = 0.0, False
    for m in materials:
        eps = np.asarray(m.eps(wl), dtype=complex)
        n_max = max(n_max, float(np.abs(np.sqrt(eps)).max()))
        metal |= bool(np.any(eps.real < 0))
    return n_max, metal


def spacing(sim, opts, feature_nm: float | None = None) -> float:
    """Cell size: spacing_nm, else the shortest wavelength over cells_per_wavelength in the densest material,
    at most the smallest particle half-extent over cells_per_radius (10, or 20 with a metal), and at most a quarter
    of ``feature_nm`` when given (scattering.py:119-133)."""
    if opts.spacing_nm:
        return float(opts.spacing_nm)
    wl = sim.wavelength_array()
    n_max, metal = _n_max([m for p in sim.particles for m in p.region_materials()], wl)
    n_max = max(n_max, sim.medium_index)
    cpr = opts.cells_per_radius or (20.0 if metal else 10.0)
    dx = wl.min() / (n_max * opts.cells_per_wavelength)
    if sim.particles:
        dx = min(dx, min(float(np.min(p.shape.bounding_half_extent())) for p in sim.particles) / cpr)
    return min(dx, feature_nm / 4) if feature_nm else dx


@dataclass
class Layout:
    """Grid around the particles (setup.py:572-647): node (0,0,0) at origin; TFSF box; total-field flux box 2 cells
    inside it, scattered-field box 2 cells outside; region examined by the rasteriser; CPML thickness."""

    dx: float
    shape: tuple
    origin: np.ndarray
    center: np.ndarray
    tfsf: tuple
    tot: tuple
    sca: tuple
    roi: tuple
    pml: int

    @property
    def cells(self) -> int:
        return int(np.prod(self.shape))


def layout(sim, opts, dx: float, bbox=None) -> Layout:
    """setup.py:627-647: odd shape centred on the particles' bounding box (or ``bbox`` = (lo, hi)); padding and gap
    4 cells, at least 3 % of the longest wavelength in the host; CPML 10 cells by default."""
    if bbox is not None:
        lo, hi = (np.asarray(v, dtype=float).copy() for v in bbox)
    else:
        lo = np.min([np.asarray(p.position_nm) - p.shape.bounding_half_extent() for p in sim.particles], axis=0)
        hi = np.max([np.asarray(p.position_nm) + p.shape.bounding_half_extent() for p in sim.particles], axis=0)
    if getattr(sim, "substrate", None) is not None:  # the box reaches the interface (scattering.py:93-99)
        lo[2], hi[2] = min(lo[2], sim.substrate.interface_z_nm), max(hi[2], sim.substrate.interface_z_nm)
    ctr = (lo + hi) / 2
    half_cells = np.ceil((hi - lo) / 2 / dx - 1e-9).astype(int) + 1
    floor = 0.03 * float(np.max(sim.wavelength_array())) / sim.medium_index
    n_pad = max(4, math.ceil((floor if opts.padding_nm is None else opts.padding_nm) / dx - 1e-9))
    n_gap = max(4, math.ceil((floor if opts.pml_gap_nm is None else opts.pml_gap_nm) / dx - 1e-9))
    half_tfsf = half_cells + n_pad
    n_pml = opts.pml_cells or 10
    half_tot = half_tfsf + 2 + n_gap + n_pml + 1
    shape = tuple(int(v) for v in 2 * half_tot + 1)
    c = half_tot.astype(int)
    roi = (np.maximum(c - half_cells - 2, 1), np.minimum(c + half_cells + 3, np.asarray(shape) - 1))
    return Layout(dx=float(dx), shape=shape, origin=ctr - c * dx, center=c, tfsf=(c - half_tfsf, c + half_tfsf),
                  tot=(c - half_tfsf + 2, c + half_tfsf - 2), sca=(c - half_tfsf - 2, c + half_tfsf + 2), roi=roi,
                  pml=n_pml)


def plan(sim, opts, media=None, warnings=None) -> dict:
    """Everything about the run that is cheap to compute (scattering.py:232-280)."""
    warnings = [] if warnings is None else warnings
    if media is None:
        media, _ = scene_media(sim, warnings)
    dx = spacing(sim, opts)
    lay = layout(sim, opts, dx)
    dt = time_step(dx, opts.courant, media)
    omega_c, omega_half, omega_max = band(sim)
    # sampling every `stride` steps aliases content above omega_c + 6 omega_half onto the band's top, never into it
    # (the table-fed wave is compared with its exact background: the e^-36 margin of 6 omega_half)
    stride = opts.dft_stride or max(1, int(2 * np.pi / (omega_c + 6 * omega_half + omega_max) / dt))
    max_t = (opts.run_time_fs or opts.max_time_fs) * FS
    max_steps = int(math.ceil(max_t / dt / stride)) * stride
    tau = 2 / omega_half
    t_pulse = 7 * tau + lay.shape[0] * dx * sim.medium_index
    gammas = [p.gamma for m in media for p in m.poles if p.sigma > 0]
    t_ring = math.log(1 / opts.decay_by) / max(min(gammas), 1e-6) if gammas else 4 * t_pulse / 7
    est_steps = max_steps if opts.run_time_fs else min(max_steps, int((t_pulse + t_ring) / dt))
    n_flux = 2 * 12 * int(np.prod(lay.sca[1] - lay.sca[0] + 1) ** (2 / 3))
    n_f = len(sim.wavelength_array())
    ext = lay.tfsf[1] - lay.tfsf[0]
    n_k = int(ext.max()) + 5
    n_bg = max(math.sqrt(media[0].eps_inf), math.sqrt(media[-1].eps_inf) if sim.substrate is not None else 0.0)
    t_tab = 2 * max(5 * tau, 30 / (2 * np.pi / float(np.max(band_wavelengths(sim))))) + 4 * tau + \
        (4 * n_k + 2 * float(np.linalg.norm(ext))) * dx * n_bg
    n_t = 1 << int(np.ceil(np.log2(t_tab / dt)))
    m = max(1, int(OMEGA_H / ((omega_c + 3 * omega_half) * dt)))
    mem = 4 * (12 * lay.cells + 4 * n_flux * n_f) + 4 * 6 * n_k * (n_t // m + 8)
    return {"dx": dx, "layout": lay, "dt": dt, "stride": stride, "max_steps": max_steps, "est_steps": est_steps,
            "omega_c": omega_c, "omega_half": omega_half, "media": media, "memory_bytes": int(mem),
            "warnings": warnings}


# --- solve ---------------------------------------------------------------------------------------------------------

@dataclass
class Scene:
    """The host inputs of one scattering run (what the engine is built from): layout, time step, DFT stride, step
    cap, material arrays and groups, CPML, the table-fed source and the monitors' index sets."""

    media: list
    layout: Layout
    dt: float
    stride: int
    max_steps: int
    inv_eps: np.ndarray
    groups: list
    offdiag: object
    pml: PMLSpec
    source: ObliqueLayeredPlaneWave
    omega: np.ndarray
    sca: tuple
    tot: tuple
    stats: dict
    mode: str
    warnings: list
    voxelise_s: float
    substrate: dict | None = None
    omega_ff: np.ndarray | None = None  # the far field's frequencies, if it is asked for
    near: dict | None = None  # the near field's slab of nodes and frequencies, if it is asked for
    probes: np.ndarray | None = None  # E nodes sampled for the resonances, if they are asked for

    def monitors(self):
        """(scattered-field box, total-field box, incident reference) DFT monitors, energy groups 1 and 2, then,
        with a far field, the scattered-field box and the incident reference at its frequencies."""
        out = [DFTMonitor(self.omega, self.sca[0], self.sca[1], self.sca[2], energy=1),
               DFTMonitor(self.omega, self.tot[0], self.tot[1], self.tot[2], energy=2),
               DFTMonitor(self.omega, np.arange(3), source=self.source.view)]
        if self.omega_ff is not None:
            out += [DFTMonitor(self.omega_ff, self.sca[0], self.sca[1], self.sca[2]),
                    DFTMonitor(self.omega_ff, np.arange(3), source=self.source.view)]
        if self.near is not None:
            out += [DFTMonitor(self.near["omega"], self.near["idx"]),
                    DFTMonitor(self.near["omega"], np.arange(3), source=self.source.view)]
        return out

    def engine(self, device, **kw) -> Engine:
        """The HIP engine of the scene, with fresh monitors (and a time monitor on the resonance probes)."""
        tm = () if self.probes is None else (TimeMonitor(self.probes, self.max_steps // self.stride + 4),)
        eng = Engine(device, shape=self.layout.shape, dx=self.layout.dx, dt=self.dt, single=True,
                     inv_eps=self.inv_eps, groups=self.groups, pml=self.pml, sources=[self.source],
                     monitors=self.monitors(), dft_stride=self.stride, offdiag=self.offdiag, time_monitors=tm, **kw)
        eng.set_tiling(*filling_tiling([eng.shape]))
        return eng


def _raster_device(device):
    """The torch device the rasteriser classifies points on: the engine's GPU, or the CPU for host-only set-up."""
    if device is None:
        return torch.device("cpu")
    if hasattr(device, "framework") and hasattr(device, "name"):
        return torch.device(device.name)
    return torch.device(device)


def setup(sim, opts, device) -> Scene:
    """Plan, media, voxelised materials (on the device), CPML, table-fed source and flux boxes of ``sim``
    (scattering.py:301-390)."""
    warnings: list[str] = []
    wl = sim.wavelength_array()
    media, slot_of = scene_media(sim, warnings)
    p = plan(sim, opts, media, warnings)
    lay, dx, dt, stride = p["layout"], p["dx"], p["dt"], p["stride"]
    if lay.cells > opts.max_cells:
        raise ValueError(f"grid {lay.shape} has {lay.cells} cells > max_cells={opts.max_cells}; "
                         "increase spacing_nm or raise max_cells")
    khat, ehat = sim.illumination.unit_vectors()
    eps_b = sim.medium_index ** 2
    mode = opts.mesh_refinement or ("conformal" if opts.subpixel else "staircase")
    slots = ParticleSlots(sim.particles, slot_of)
    layered, sub_kw, substrate = {}, {}, None
    if sim.substrate is not None:
        z_int, s_sub = float(sim.substrate.interface_z_nm), len(media) - 1
        slots = OnSubstrate(slots, z_int, s_sub)
        layered = dict(base=planar_background(lay.shape, lay.origin, dx, z_int, media[0], media[s_sub]),
                       background=(0, s_sub))
        sub_kw = dict(z_int=z_int, sub=media[s_sub])
        substrate = {"interface_z_nm": z_int, "incidence": "from above" if khat[2] < 0 else "through the substrate",
                     "source": "layered tables", "absorbing": False}
    t0 = time.perf_counter()
    inv_eps, groups, vstats, offdiag = voxelise(slots, media, lay.shape, lay.origin, dx, lay.roi[0], lay.roi[1],
                                                mode=mode, samples=opts.subpixel_samples,
                                                device=_raster_device(device), **layered)
    t_vox = time.perf_counter() - t0
    lam_max = float(wl.max())
    pml = PMLSpec(cells=((lay.pml, lay.pml),) * 3, kappa_max=opts.pml_kappa,
                  alpha_max=0.05 * 2 * np.pi * sim.medium_index / lam_max, eps_ref=eps_b)
    src = ObliqueLayeredPlaneWave(shape=lay.shape, origin=lay.origin, spacing=dx, dt=dt, lo=lay.tfsf[0],
                                  hi=lay.tfsf[1], khat=khat, ehat=ehat, inv_eps=inv_eps, host=media[0],
                                  omega_c=p["omega_c"], omega_half=p["omega_half"],
                                  max_steps=p["max_steps"] + 4 * stride,
                                  omega_min=2 * np.pi / float(np.max(band_wavelengths(sim))), **sub_kw)
    return Scene(media=media, layout=lay, dt=dt, stride=stride, max_steps=p["max_steps"], inv_eps=inv_eps,
                 groups=groups, offdiag=offdiag, pml=pml, source=src, omega=2 * np.pi / wl,
                 sca=flux_points(*lay.sca, lay.shape, dx), tot=flux_points(*lay.tot, lay.shape, dx), stats=vstats,
                 mode=mode, warnings=warnings, voxelise_s=t_vox, substrate=substrate,
                 omega_ff=None if sim.outputs.far_field is None else
                 2 * np.pi / np.asarray(sim.outputs.far_field.wavelengths_nm, dtype=float),
                 near=None if sim.outputs.near_field is None else near_slab(sim, lay),
                 probes=None if sim.outputs.resonances is None else resonance_probes(sim, lay))


def cross_sections(scene: Scene, e_ref, es, hs, et, ht) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
    """(c_sca, c_abs, c_ext) from the monitors' spectra (scattering.py:428-464): incident intensity at the
    reference point n |E|^2 / 2; the background's own flux through the total-field box from the same flux formula."""
    src, omega, tot = scene.source, scene.omega, scene.tot
    intensity = 0.5 * src.n_inc * np.sum(np.abs(e_ref) ** 2, axis=1)
    unit = src.incident_reference(omega)
    amp = np.sum(e_ref * np.conj(unit), axis=1) / np.sum(np.abs(unit) ** 2, axis=1)
    e_bg = amp[:, None] * src.background(omega, tot[0], "E")
    h_bg = amp[:, None] * 0.5 * (src.background(omega, tot[1], "H") + src.background(omega, tot[2], "H"))
    bg_out = flux(e_bg, h_bg, tot[3])
    c_sca = flux(es, hs, scene.sca[3]) / intensity
    c_abs = -(flux(et, ht, tot[3]) - bg_out) / intensity
    return c_sca, c_abs, c_sca + c_abs


def far_field_data(scene: Scene, sim, e_ff, h_ff, e_inc) -> list:
    """dC_sca/dOmega on the requested grid from the scattered-field box's near-to-far transform, per unit incident
    intensity (scattering.py:592-625): theta from the incident direction, phi from the incident polarisation."""
    from ..problem import FarFieldData
    req = sim.outputs.far_field
    khat, ehat = sim.illumination.unit_vectors()
    th = np.radians(np.asarray(req.theta_deg, dtype=float))[:, None, None]
    ph = np.radians(np.asarray(req.phi_deg, dtype=float))[None, :, None]
    u = np.cos(th) * khat + np.sin(th) * (np.cos(ph) * ehat + np.sin(ph) * np.cross(khat, ehat))
    lay = scene.layout
    surf = farfield.box_surface(scene.sca[0], scene.sca[1], scene.sca[3], lay.shape, lay.origin, lay.dx)
    nb = sim.medium_index
    out = []
    for f, lam in enumerate(req.wavelengths_nm):
        ef = farfield.far_field(surf, e_ff[f], h_ff[f], u.reshape(-1, 3), nb * scene.omega_ff[f], 1 / nb,
                                center=lay.origin + lay.center * lay.dx)
        intensity = 0.5 * scene.source.n_inc * float(np.sum(np.abs(e_inc[f]) ** 2))
        dc = (0.5 * nb * np.sum(np.abs(ef) ** 2, axis=1) / intensity).reshape(u.shape[:2])
        out.append(FarFieldData(wavelength_nm=float(lam), theta_deg=list(req.theta_deg), phi_deg=list(req.phi_deg),
                                dcsca_domega_nm2=dc.tolist()))
    return out


def resonance_probes(sim, lay: Layout, max_particles: int = 4) -> np.ndarray:
    """E nodes (all three components) at each of the first particles' centre and at 14 points two cells outside
    its bounding box, along the axes and the diagonals (scattering.py:506-521)."""
    dirs = np.concatenate([np.eye(3), -np.eye(3), np.array(np.meshgrid(*[(1, -1)] * 3, indexing="ij")).reshape(3, -1).T
                           / np.sqrt(3)])
    idx = []
    for p in sim.particles[:max_particles]:
        c = np.asarray(p.position_nm, dtype=float)
        h = np.asarray(p.shape.bounding_half_extent(), dtype=float)
        for pt in [c] + [c + d * (h + 2 * lay.dx) for d in dirs]:
            for comp in range(3):
                f = (pt - lay.origin) / lay.dx - 0.5 * np.eye(3)[comp]
                ijk = np.clip(np.rint(f).astype(int), 1, np.asarray(lay.shape) - 2)
                idx.append(int(flat_index(comp, *ijk, lay.shape)))
    return np.unique(np.asarray(idx, dtype=np.int64))


def resonances(sim, x: np.ndarray, dt: float, warnings: list) -> list:
    """The modes of the probes' ring-down in the band, amplitudes relative to the strongest (scattering.py:524-536)."""
    from ..problem import ResonanceData
    from .resonance import harmonic_inversion
    wl = sim.wavelength_array()
    if x.shape[1] < 64:
        warnings.append("resonances: the field decayed too fast after the pulse for harmonic inversion; no "
                        "resonance found (lower decay_by or set run_time_fs)")
        return []
    modes = harmonic_inversion(x, dt, 2 * np.pi / wl.max(), 2 * np.pi / wl.min(),
                               max_modes=sim.outputs.resonances.max_modes)
    if not modes:
        return []
    a0 = modes[0].amplitude
    return [ResonanceData(wavelength_nm=2 * np.pi / m.omega.real, energy_ev=HBARC_EV_NM * m.omega.real, q=m.q,
                          amplitude=m.amplitude / a0) for m in modes]


_PLANE_AXES = {"xy": (0, 1), "xz": (0, 2), "yz": (1, 2)}


def near_slab(sim, lay: Layout) -> dict:
    """The E nodes (all three components) of the slab of cells around the requested plane that the near field is
    interpolated from, one cell beyond it on every side (scattering.py:631-652)."""
    req = sim.outputs.near_field
    a1, a2 = _PLANE_AXES[req.plane]
    normal = 3 - a1 - a2
    ranges = []
    for a in range(3):
        half = lay.dx if a == normal else req.half_width_nm + lay.dx
        centre = req.offset_nm if a == normal else 0.0
        lo = max(1, int(math.floor((centre - half - lay.origin[a]) / lay.dx)) - 1)
        hi = min(lay.shape[a] - 2, int(math.ceil((centre + half - lay.origin[a]) / lay.dx)) + 1)
        if hi <= lo:
            raise ValueError("near_field plane lies outside the FDTD grid; reduce half_width_nm or offset_nm")
        ranges.append(np.arange(lo, hi + 1))
    gi, gj, gk = np.meshgrid(*ranges, indexing="ij")
    idx = np.concatenate([flat_index(c, gi, gj, gk, lay.shape).ravel() for c in range(3)])
    return {"ranges": ranges, "idx": idx,
            "omega": 2 * np.pi / np.asarray(req.wavelengths_nm, dtype=float)}


def near_field_data(scene: Scene, sim, e_slab, e_inc, warnings=None) -> list:
    """The total field on the requested plane (scattering.py:655-735): the slab's DFT, plus the background wave
    from the source tables at nodes outside the total-field box (where the DFT holds the scattered field),
    each component interpolated trilinearly from its own Yee positions to the points, over the incident amplitude
    at the origin."""
    from scipy.interpolate import RegularGridInterpolator

    from ..problem import NearFieldData
    req = sim.outputs.near_field
    a1, a2 = _PLANE_AXES[req.plane]
    normal = 3 - a1 - a2
    lay, src, near = scene.layout, scene.source, scene.near
    rng, om = near["ranges"], near["omega"]
    shp = tuple(len(r) for r in rng)
    e_all = np.asarray(e_slab).reshape(len(om), 3, *shp)
    u = np.linspace(-req.half_width_nm, req.half_width_nm, req.points)
    g1, g2 = np.meshgrid(u, u, indexing="ij")
    pts = np.zeros(g1.shape + (3,))
    pts[..., a1], pts[..., a2], pts[..., normal] = g1, g2, req.offset_nm
    khat, ehat = sim.illumination.unit_vectors()
    unit = src.incident_reference(om)
    amp = np.sum(e_inc * np.conj(unit), axis=1) / np.sum(np.abs(unit) ** 2, axis=1)  # along ehat at the reference
    inc0 = amp * np.exp(-1j * src.n_inc * om * (khat @ src.r_ref))  # the incident amplitude at the origin
    lo, hi = lay.tfsf
    out = []
    for f, lam in enumerate(req.wavelengths_nm):
        k = sim.medium_index * om[f]
        comps = []
        for c in range(3):
            frac = [rng[a] + (0.5 if a == c else 0.0) for a in range(3)]  # fractional indices of E_c
            axes = [lay.origin[a] + frac[a] * lay.dx for a in range(3)]
            U = np.meshgrid(*frac, indexing="ij")
            inside = np.all([(U[a] >= lo[a]) & (U[a] <= hi[a]) for a in range(3)], axis=0)
            nodes = flat_index(c, *[np.floor(U[a] + 1e-9).astype(int).ravel() for a in range(3)], lay.shape)
            bg = (amp[f] * src.background(om[f:f + 1], nodes, "E")[0]).reshape(U[0].shape)
            if sim.substrate is None:  # beyond the tables: the continuum plane wave
                P_ = np.stack(np.meshgrid(*axes, indexing="ij"), axis=-1)
                bg = np.where(np.isnan(bg), ehat[c] * inc0[f] * np.exp(1j * k * (P_ @ khat)), bg)
            vals = e_all[f, c] + np.where(inside, 0.0, bg)
            interp = RegularGridInterpolator(axes, vals, bounds_error=False, fill_value=np.nan)
            comps.append(interp(pts))
        e = np.stack(comps, axis=-1) / inc0[f]
        if warnings is not None and f == 0 and np.isnan(e).any():
            warnings.append("near_field: points outside the source box beyond the range where the background field "
                            "is known are NaN")
        out.append(NearFieldData(wavelength_nm=float(lam), plane=req.plane, offset_nm=req.offset_nm,
                                 axis1_nm=u.tolist(), axis2_nm=u.tolist(),
                                 enhancement=np.sum(np.abs(e) ** 2, axis=-1).tolist(),
                                 e_real=np.moveaxis(e.real, -1, 0).tolist(),
                                 e_imag=np.moveaxis(e.imag, -1, 0).tolist()))
    return out


def solve(sim, opts, device) -> dict:
    """Cross sections of ``sim`` (c_sca, c_abs, c_ext in nm^2 per wavelength) with provenance and warnings."""
    t_start = time.perf_counter()
    scene = setup(sim, opts, device)
    warnings = scene.warnings
    lay, dt, stride = scene.layout, scene.dt, scene.stride
    eng = scene.engine(device)
    mon_sca, mon_tot, mon_inc = eng.monitors[:3]
    far = near = res_data = None
    t_loop = time.perf_counter()
    try:
        steps, level, graphs = run_until_decayed(eng, opts, max_steps=scene.max_steps,
                                                 min_steps=int(scene.source.t_passed / dt),
                                                 window=math.ceil(float(2 * np.pi / scene.omega.min()) / (stride * dt)),
                                                 warnings=warnings)
        t_loop = time.perf_counter() - t_loop
        e_ref, _ = mon_inc.results()
        es, hs = mon_sca.results()
        et, ht = mon_tot.results()
        nxt = 3
        if scene.omega_ff is not None:
            e_ff, h_ff = eng.monitors[nxt].results()
            far = far_field_data(scene, sim, e_ff, h_ff, eng.monitors[nxt + 1].results()[0])
            nxt += 2
        if scene.near is not None:
            near = near_field_data(scene, sim, eng.monitors[nxt].results()[0], eng.monitors[nxt + 1].results()[0],
                                   warnings)
        if scene.probes is not None:  # the ring-down once the source is off
            n_skip = int(math.ceil(scene.source.t_off / (stride * dt)))
            res_data = resonances(sim, eng.time_monitors[0].results(steps // stride)[:, n_skip:], stride * dt,
                                  warnings)
    finally:
        eng.close()
    c_sca, c_abs, c_ext = cross_sections(scene, e_ref, es, hs, et, ht)
    if np.any(c_abs < -1e-3 * np.max(np.abs(c_ext))):
        warnings.append("negative absorption beyond 0.1% of extinction: increase resolution or decay time")
    prov = {"device": str(eng._dev), "grid": list(lay.shape), "cells": lay.cells, "spacing_nm": lay.dx,
            "dt_fs": dt / FS, "steps": steps, "simulated_time_fs": steps * dt / FS, "dft_stride": stride,
            "precision": "single", "hip_graphs": graphs, "subpixel": scene.mode != "staircase", "pml_cells": lay.pml,
            "tfsf_box_nm": float((lay.tfsf[1] - lay.tfsf[0])[0] * lay.dx), "decay_level": level,
            "materials": [m.describe() for m in scene.media], **scene.stats, "symmetry": "off",
            "tfsf_table_stride": scene.source.m, "source": "table-fed plane wave",
            "runtime_s": time.perf_counter() - t_start, "time_loop_s": t_loop, "voxelise_s": scene.voxelise_s,
            **({"substrate": scene.substrate} if scene.substrate is not None else {})}
    return {"c_sca": c_sca, "c_abs": c_abs, "c_ext": c_ext, "near_field": near, "far_field": far,
            "resonances": res_data, "provenance": prov, "warnings": warnings}


__all__ = ["Layout", "Scene", "band", "band_wavelengths", "cross_sections", "far_field_data", "layout",
           "near_field_data", "near_slab", "particle_media", "plan", "resonance_probes", "resonances", "scene_media",
           "setup", "solve", "spacing", "substrate_unsupported", "to_medium"]
