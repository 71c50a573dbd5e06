"""Problem-level drivers on the native engine: plane-wave scattering by particles, periodic reflection and
transmission per diffraction order, and dipole emission (scattering.py:232-500; periodic.py:392-459, 503-519;
emission.py:153-196).

Each driver assembles the engine parts for its problem kind -- layout, time step and DFT stride from the source
band, material arrays, CPML, source, flux monitors -- runs them under ``run_until_decayed`` and post-processes the
monitors on the host, with the reference's formulas:

* ``solve_scattering``: TFSF plane wave around the particles; C_sca from the outward flux through the box in the
  scattered-field region, C_abs from the inward flux through the box in the total-field region, both per unit
  incident intensity; C_ext = C_sca + C_abs (scattering.py:428-464).
* ``solve_periodic``: one Bloch-periodic cell, normal incidence from a 1D-fed TFSF plane or oblique incidence at
  fixed k_par from a Bloch plane source; reflection and transmission per order from the FFT of the plane monitors
  over the cell (periodic.py:347-363, 392-459). ``periodic_run`` builds such a run without stepping it and
  ``periodic_finish`` reads it out, so that ``spectra.solve`` can step many as one batch.
* ``solve_emission``: a point dipole; the power it gives to the field, -1/2 dx^3 Re(E_src . J*), and the power
  through a closed box around it (emission.py:153-196).

Scope of this module: a uniform grid and a homogeneous, lossless host; particles are given as voxelised permittivity
(``staircase`` places spheres on the grid the way the reference's mesh_refinement="staircase" does: the medium at
each E node's position). The plane wave of ``solve_scattering`` comes from the 1D-fed TFSF source
(tfsf.py:108-211), which the reference uses for periodic normal incidence. The problem-level solves with the
reference's schemas sit above these drivers: ``scattering.solve`` (table-fed TFSF from the exact discrete
background, ``layered.py``) and ``spectra.solve`` (layer stacks and periodic structures).
"""

from __future__ import annotations

import math
import time
from dataclasses import dataclass, field

import numpy as np

from .engine import DFTMonitor, DispersiveGroup, Engine, PMLSpec, filling_tiling, flat_index
from .media import Medium, limit_dt, time_step
from .run import FS, RunOptions, flux, flux_points, orders, plane_points, run_until_decayed
from .sources import BlochPlaneSource, PointDipole, TFSFPlaneWave


@dataclass
class DriverOptions(RunOptions):
    """The FDTDOptions fields the drivers read (options.py:28-92), with the reference's defaults."""

    courant: float = 0.95
    pml_cells: int | None = None
    pml_kappa: float = 1.0
    padding_nm: float | None = None
    pml_gap_nm: float | None = None
    dft_stride: int | None = None


def band(wavelengths) -> tuple[float, float, float]:
    """(omega_c, omega_half, omega_max) of the source pulse that covers the wavelengths (scattering.py:143-152)."""
    om = 2 * np.pi / np.asarray(wavelengths, dtype=float)
    omega_c = 0.5 * (om.min() + om.max())
    omega_half = max(0.5 * (om.max() - om.min()), 0.15 * omega_c)
    return float(omega_c), float(omega_half), float(om.max())


# --- layout and voxelisation ----------------------------------------------------------------------------------

@dataclass
class Layout:
    """A scattering grid (setup.py:592-647): node (0,0,0) at `origin`; index boxes of the TFSF surface, the
    total-field flux box (inside it) and the scattered-field flux box (outside it); CPML thickness."""

    dx: float
    shape: tuple
    origin: np.ndarray
    center: np.ndarray
    tfsf: tuple
    tot: tuple
    sca: tuple
    pml: int

    @property
    def cells(self) -> int:
        return int(np.prod(self.shape))


def scattering_layout(bbox_lo, bbox_hi, dx: float, wl_max: float, n_host: float = 1.0,
                      opts: DriverOptions | None = None) -> Layout:
    """Grid around the particles' bounding box (setup.py:627-647): padding to the TFSF box and gap to the CPML
    of 4 cells, at least 3 % of the longest wavelength in the host; flux boxes 2 cells inside and outside the
    TFSF box; odd shape, centred on the box."""
    opts = opts or DriverOptions()
    lo, hi = np.asarray(bbox_lo, float), np.asarray(bbox_hi, float)
    ctr = (lo + hi) / 2
    half_cells = np.ceil((hi - lo) / 2 / dx - 1e-9).astype(int) + 1
    floor = 0.03 * wl_max / n_host
    n_pad = max(4, math.ceil((floor if opts.padding_nm is None else opts.padding_nm) / dx - 1e-9))
    n_gap = max(4, math.ceil((floor if opts.pml_gap_nm is None else opts.pml_gap_nm) / dx - 1e-9))
    n_pml = opts.pml_cells or 10
    half_tfsf = half_cells + n_pad
    half_tot = half_tfsf + 2 + n_gap + n_pml + 1
    shape = tuple(int(v) for v in 2 * half_tot + 1)
    c = half_tot.astype(int)
    return Layout(dx=float(dx), shape=shape, origin=ctr - c * dx, center=c, tfsf=(c - half_tfsf, c + half_tfsf),
                  tot=(c - half_tfsf + 2, c + half_tfsf - 2), sca=(c - half_tfsf - 2, c + half_tfsf + 2), pml=n_pml)


@dataclass
class Sphere:
    center: tuple
    radius: float
    medium: Medium


def e_node_positions(shape, origin, dx: float, c: int):
    """World coordinates (3 arrays over the grid) of the E_c nodes: node + e_c / 2."""
    axes = [origin[a] + (np.arange(shape[a]) + (0.5 if a == c else 0.0)) * dx for a in range(3)]
    return np.meshgrid(*axes, indexing="ij")


def staircase(shape, origin, dx: float, host: Medium, spheres) -> tuple[np.ndarray, list]:
    """(inv_eps, groups): every E node takes the medium at its position (the last sphere that contains it wins).
    A non-dispersive medium sets 1/eps; a dispersive one makes the node a D-form node (1/eps = 1) of that
    medium's pure group (setup.py:187, 234; engine.py:722-742)."""
    if host.poles:
        raise ValueError("the host medium must be non-dispersive")
    inv_eps = np.full((3,) + tuple(shape), 1.0 / host.eps_inf)
    owner = np.full((3,) + tuple(shape), -1, dtype=np.int32)
    for c in range(3):
        X, Y, Z = e_node_positions(shape, np.asarray(origin, float), dx, c)
        for s, sp in enumerate(spheres):
            cx, cy, cz = sp.center
            inside = (X - cx) ** 2 + (Y - cy) ** 2 + (Z - cz) ** 2 <= sp.radius ** 2
            owner[c][inside] = s
    groups = []
    by_medium: dict[int, list] = {}
    for s, sp in enumerate(spheres):
        m = sp.medium
        sel = owner == s
        if m.poles:
            by_medium.setdefault(id(m), [m, []])[1].append(np.nonzero(sel.reshape(-1))[0])
            inv_eps[sel] = 1.0
        else:
            inv_eps[sel] = 1.0 / m.eps_inf
    for m, parts in by_medium.values():
        idx = np.sort(np.concatenate(parts)).astype(np.int64)
        if len(idx):
            groups.append(DispersiveGroup(idx=idx, media=[m], frac=np.ones((1, len(idx))), w=None))
    return inv_eps, groups


# --- scattering ------------------------------------------------------------------------------------------------

@dataclass
class ScatteringResult:
    wavelengths: np.ndarray
    c_sca: np.ndarray
    c_abs: np.ndarray
    c_ext: np.ndarray
    provenance: dict = field(default_factory=dict)
    warnings: list = field(default_factory=list)


def plan_scattering(layout: Layout, media, wavelengths, opts: DriverOptions, n_host: float = 1.0) -> dict:
    """dt, DFT stride, step cap and source band of a scattering run (scattering.py:232-280): the stride keeps
    source content above omega_c + 4 omega_half from aliasing onto the band."""
    dt = time_step(layout.dx, opts.courant, media)
    omega_c, omega_half, omega_max = band(wavelengths)
    stride = opts.dft_stride or max(1, int(2 * np.pi / (omega_c + 4 * omega_half + omega_max) / dt))
    max_t = (opts.run_time_fs or opts.max_time_fs) * FS
    max_steps = int(math.ceil(max_t / dt / stride)) * stride
    return {"dt": dt, "stride": stride, "max_steps": max_steps, "omega_c": omega_c, "omega_half": omega_half}


def solve_scattering(device, layout: Layout, inv_eps, groups, media, wavelengths, khat, ehat, *, n_host: float = 1.0,
                     opts: DriverOptions | None = None, offdiag=None) -> ScatteringResult:
    """Cross sections (same length unit as dx, squared) per wavelength of the particles voxelised into
    inv_eps / groups, for a plane wave along khat polarised along ehat in a host of index n_host."""
    t_start = time.perf_counter()
    opts = opts or DriverOptions()
    warnings: list[str] = []
    wl = np.asarray(wavelengths, dtype=float)
    p = plan_scattering(layout, media, wl, opts, n_host)
    dt, stride, dx = p["dt"], p["stride"], layout.dx
    eps_b = n_host ** 2
    khat, ehat = np.asarray(khat, float), np.asarray(ehat, float)
    khat = khat / np.linalg.norm(khat)
    ehat = ehat - (ehat @ khat) * khat
    ehat = ehat / np.linalg.norm(ehat)
    pml = PMLSpec(cells=((layout.pml, layout.pml),) * 3, kappa_max=opts.pml_kappa,
                  alpha_max=0.05 * 2 * np.pi * n_host / float(wl.max()), eps_ref=eps_b)
    r0 = layout.origin + layout.center * dx
    src = TFSFPlaneWave(shape=layout.shape, origin=layout.origin, dx=dx, dt=dt, r0=r0, lo=layout.tfsf[0],
                        hi=layout.tfsf[1], khat=khat, ehat=ehat, eps_b=eps_b, omega_c=p["omega_c"],
                        omega_half=p["omega_half"], max_steps=p["max_steps"] + 4 * stride)
    omega = 2 * np.pi / wl
    sca = flux_points(*layout.sca, layout.shape, dx)
    tot = flux_points(*layout.tot, layout.shape, dx)
    mon_sca = DFTMonitor(omega, sca[0], sca[1], sca[2], energy=1)
    mon_tot = DFTMonitor(omega, tot[0], tot[1], tot[2], energy=2)
    mon_inc = DFTMonitor(omega, *src.probe(), source=src)
    eng = Engine(device, shape=layout.shape, dx=dx, dt=dt, single=True, inv_eps=inv_eps, groups=groups, pml=pml,
                 sources=[src], monitors=[mon_sca, mon_tot, mon_inc], dft_stride=stride, offdiag=offdiag)
    t_loop = time.perf_counter()
    try:
        steps, level, graphs = run_until_decayed(eng, opts, max_steps=p["max_steps"], min_steps=int(src.t_passed / dt),
                                                 window=math.ceil(float(wl.max()) / (stride * dt)), warnings=warnings)
        t_loop = time.perf_counter() - t_loop
        e_i, h_i = mon_inc.results()
        es, hs = mon_sca.results()
        et, ht = mon_tot.results()
    finally:
        eng.close()
    # incident intensity on the line at the TFSF reference point: 0.5 Re(E H*) with H averaged to the E node
    intensity = np.abs(0.5 * np.real(e_i[:, 0] * np.conj(h_i[:, 0])))
    c_sca = flux(es, hs, sca[3]) / intensity
    c_abs = -flux(et, ht, tot[3]) / intensity  # a lossless homogeneous background carries no net flux
    c_ext = c_sca + c_abs
    if np.any(c_abs < -1e-3 * np.max(np.abs(c_ext))):
        warnings.append("negative absorption beyond 0.1% of extinction: increase resolution or decay time")
    prov = {"grid": list(layout.shape), "cells": layout.cells, "spacing_nm": dx, "dt_fs": dt / FS, "steps": steps,
            "dft_stride": stride, "decay_level": level, "hip_graphs": graphs, "pml_cells": layout.pml,
            "source": "1D-fed TFSF", "time_loop_s": t_loop, "runtime_s": time.perf_counter() - t_start}
    return ScatteringResult(wl, c_sca, c_abs, c_ext, prov, warnings)


# --- periodic cells --------------------------------------------------------------------------------------------

@dataclass
class PeriodicCell:
    """A Bloch-periodic cell of n_cells real cells per lateral axis (the engine adds one ghost plane per side),
    nz nodes along z; the source plane k_src, the reflection monitor plane k_r above it, the transmission plane k_t
    below the structure, the incident reference plane k_top (periodic.py:145-190)."""

    n_cells: tuple
    nz: int
    spacing: tuple
    inv_eps: np.ndarray
    groups: list
    pml: int
    k_src: int
    k_r: int
    k_t: int
    k_top: int
    eps_sup: float = 1.0
    eps_sub: float = 1.0
    offdiag: object = None  # engine.OffDiagonal of conformal dielectric interfaces, or None
    origin: np.ndarray | None = None  # nm; None: the grid's first node at 0

    @property
    def shape(self) -> tuple:
        return (self.n_cells[0] + 2, self.n_cells[1] + 2, self.nz)


def periodic_time_step(cell: PeriodicCell, courant: float, media, kpar=(0.0, 0.0)) -> float:
    """Courant step of a periodic cell (periodic.py:305-321): a one-cell periodic axis carries the single Bloch
    wavenumber k, grid eigenvalue 2 sin(k d / 2) / d; scaled by sqrt(min eps_inf) and lowered where an ADE medium
    would be unstable (setup.py:72-75)."""
    d = cell.spacing
    inv2 = 1 / d[2] ** 2
    for a in (0, 1):
        if cell.n_cells[a] > 1:
            inv2 += 1 / d[a] ** 2
        elif kpar[a]:
            inv2 += (math.sin(kpar[a] * d[a] / 2) / d[a]) ** 2
    dt0 = courant / math.sqrt(inv2)
    scale = math.sqrt(min([1.0] + [m.eps_inf for m in media]))  # periodic.py:479 (no faster than in vacuum)
    return limit_dt(dt0 * scale, courant, media, 4 * (courant / dt0) ** 2)


@dataclass
class PeriodicResult:
    r: np.ndarray  # (n_f, nx, ny) reflected power per order, FFT bin layout
    t: np.ndarray
    mx: np.ndarray
    my: np.ndarray
    provenance: dict = field(default_factory=dict)
    warnings: list = field(default_factory=list)


@dataclass
class PeriodicRun:
    """One run of a periodic cell, built and not yet stepped: its engine, monitors and run-control limits."""

    eng: Engine
    src: object
    mon_r: DFTMonitor
    mon_t: DFTMonitor
    mon_i: DFTMonitor
    pol: str
    kpar: tuple
    oblique: bool
    stride: int
    max_steps: int
    min_steps: int
    window: int
    area: float
    n_cells: tuple
    spacing: tuple
    dt: float


def periodic_plan(cell: PeriodicCell, wavelengths, kpar=(0.0, 0.0), opts: DriverOptions | None = None,
                  media_extra=()) -> dict:
    """dt, DFT stride and pulse band of one run (periodic.py:275-331): normal incidence one broadband pulse over the
    band, oblique incidence a pulse aimed at the first wavelength and kept clear of the grazing frequency."""
    opts = opts or DriverOptions()
    wl = np.atleast_1d(np.asarray(wavelengths, dtype=float))
    oblique = any(abs(k) > 0 for k in kpar)
    media = [Medium(cell.eps_sup), Medium(cell.eps_sub)] + [m for g in cell.groups for m in g.media] + list(media_extra)
    dt = periodic_time_step(cell, opts.courant, media, kpar if oblique else (0.0, 0.0))
    if not oblique:
        omega_c, omega_half, omega_max = band(wl)
        stride = max(1, int(2 * np.pi / (omega_c + 4 * omega_half + omega_max) / dt))
    else:
        omega_c = 2 * np.pi / wl[0]
        k = math.hypot(*kpar)
        gaps = [abs(omega_c - k / math.sqrt(e)) for e in (cell.eps_sup, cell.eps_sub) if e > 0]
        omega_half = float(np.clip(min(gaps) / 4, 0.01 * omega_c, 0.15 * omega_c))  # periodic.py:324-331
        stride = max(1, int(2 * np.pi / (2 * omega_c + 4 * omega_half) / dt))
    return {"dt": dt, "stride": opts.dft_stride or stride, "omega_c": omega_c, "omega_half": omega_half,
            "media": media}


def periodic_run(device, cell: PeriodicCell, wavelengths, *, pol: str = "s", kpar=(0.0, 0.0), phi_deg: float = 0.0,
                 opts: DriverOptions | None = None, stride: int | None = None, media_extra=()) -> PeriodicRun:
    """The engine of one run (periodic.py:392-459): kpar = 0 one broadband real run fed by a 1D TFSF plane;
    otherwise one complex run at the first wavelength with a Bloch plane source. ``stride`` overrides the run's
    own DFT stride (a batch of runs shares one; a smaller stride than the run's own is always valid)."""
    opts = opts or DriverOptions()
    wl = np.atleast_1d(np.asarray(wavelengths, dtype=float))
    d = tuple(float(v) for v in cell.spacing)
    shape = cell.shape
    lam_max = float(wl.max())
    oblique = any(abs(k) > 0 for k in kpar)
    pml = PMLSpec(cells=((0, 0), (0, 0), (cell.pml, cell.pml)), kappa_max=opts.pml_kappa,
                  alpha_max=0.05 * 2 * np.pi * math.sqrt(cell.eps_sup) / lam_max, eps_ref=cell.eps_sup,
                  eps_ref_high=cell.eps_sub)
    p = periodic_plan(cell, wl, kpar, opts, media_extra)
    dt, stride = p["dt"], stride or p["stride"]
    media = p["media"]
    transit = shape[2] * d[2] * max(math.sqrt(abs(m.eps_inf)) for m in media)
    max_steps = int(math.ceil((opts.run_time_fs or opts.max_time_fs) * FS / dt / stride)) * stride
    origin = np.zeros(3) if cell.origin is None else np.asarray(cell.origin, dtype=float)
    if not oblique:
        omega = 2 * np.pi / wl
        phi = math.radians(phi_deg)
        ehat = np.array([-np.sin(phi), np.cos(phi), 0.0]) if pol == "s" else np.array([np.cos(phi), np.sin(phi), 0.0])
        r0 = origin + np.array([1, 1, cell.k_top]) * np.asarray(d)
        big = 1 << 30
        src = TFSFPlaneWave(shape=shape, origin=origin, dx=d, dt=dt, r0=r0, lo=np.array([-big, -big, cell.k_src]),
                            hi=np.array([big, big, big]), khat=np.array([0.0, 0.0, 1.0]), ehat=ehat,
                            eps_b=cell.eps_sup, omega_c=p["omega_c"], omega_half=p["omega_half"],
                            max_steps=max_steps + 4 * stride, periodic=(True, True, False))
        probe = src.probe()
        bloch = (0.0, 0.0, 0.0)
    else:
        omega = np.array([2 * np.pi / wl[0]])
        k = math.hypot(*kpar)
        # the column is all superstrate: its CPML is graded for that medium at both ends
        col_pml = PMLSpec(cells=pml.cells, kappa_max=pml.kappa_max, alpha_max=pml.alpha_max, eps_ref=cell.eps_sup)
        src = BlochPlaneSource(shape=shape, spacing=d, dt=dt, k_src=cell.k_src, k_launch=cell.k_src - 1, kpar=kpar,
                               pol=pol, eps_b=cell.eps_sup, pml=col_pml, omega_c=float(omega[0]),
                               omega_half=p["omega_half"], max_steps=max_steps + 4 * stride)
        probe = src.probe(cell.k_top)
        bloch = (kpar[0], kpar[1], 0.0)
        cos_t = math.sqrt(max(1 - k ** 2 / (cell.eps_sup * omega[0] ** 2), 0.01))  # slower transit along z
        transit = transit / cos_t
    mon_r = DFTMonitor(omega, *plane_points(shape, cell.n_cells, cell.k_r), energy=1)
    mon_t = DFTMonitor(omega, *plane_points(shape, cell.n_cells, cell.k_t), energy=1)
    mon_i = DFTMonitor(omega, *probe, source=src)
    eng = Engine(device, shape=shape, dx=d, dt=dt, single=True, inv_eps=cell.inv_eps, groups=cell.groups, pml=pml,
                 sources=[src], monitors=[mon_r, mon_t, mon_i], dft_stride=stride, periodic=(True, True, False),
                 bloch=bloch, offdiag=cell.offdiag)
    eng.set_tiling(*filling_tiling([shape]))
    return PeriodicRun(eng=eng, src=src, mon_r=mon_r, mon_t=mon_t, mon_i=mon_i, pol=pol,
                       kpar=tuple(kpar) if oblique else (0.0, 0.0), oblique=oblique, stride=stride,
                       max_steps=max_steps, min_steps=int((src.t_passed + transit) / dt),
                       window=math.ceil((lam_max if not oblique else float(wl[0])) / (stride * dt)),
                       area=cell.n_cells[0] * d[0] * cell.n_cells[1] * d[1], n_cells=tuple(cell.n_cells), spacing=d,
                       dt=dt)


def periodic_finish(run: PeriodicRun) -> tuple[np.ndarray, np.ndarray, np.ndarray, np.ndarray]:
    """(r, t per order (n_f, nx, ny), mx, my) of a stepped run: plane fluxes split into orders by the FFT over the
    cell (periodic.py:347-363), over the incident power through the cell (periodic.py:421-422, 455-456)."""
    e_i, h_i = run.mon_i.results()
    er, hr = run.mon_r.results()
    et, ht = run.mon_t.results()
    if run.oblique:
        inc = 0.5 * np.real(e_i[:, 0] * np.conj(h_i[:, 0]) - e_i[:, 1] * np.conj(h_i[:, 1])) * run.area
    else:
        inc = 0.5 * np.real(e_i[:, 0] * np.conj(h_i[:, 0])) * run.area
    fr, mx, my = orders(er, hr, run.n_cells, run.spacing, run.kpar)
    ft, _, _ = orders(et, ht, run.n_cells, run.spacing, run.kpar)
    return -fr / inc[:, None, None], ft / inc[:, None, None], mx, my


def solve_periodic(device, cell: PeriodicCell, wavelengths, *, pol: str = "s", kpar=(0.0, 0.0), phi_deg: float = 0.0,
                   opts: DriverOptions | None = None) -> PeriodicResult:
    """Reflection and transmission per diffraction order of one run (periodic.py:392-459), stepped alone under
    ``run_until_decayed``."""
    opts = opts or DriverOptions()
    warnings: list[str] = []
    run = periodic_run(device, cell, wavelengths, pol=pol, kpar=kpar, phi_deg=phi_deg, opts=opts)
    t_loop = time.perf_counter()
    try:
        steps, level, graphs = run_until_decayed(run.eng, opts, max_steps=run.max_steps, min_steps=run.min_steps,
                                                 window=run.window, warnings=warnings)
        t_loop = time.perf_counter() - t_loop
        r, t, mx, my = periodic_finish(run)
    finally:
        run.eng.close()
    prov = {"polarization": pol, "kpar": list(run.kpar), "steps": steps, "decay_level": level, "hip_graphs": graphs,
            "dft_stride": run.stride, "dt_fs": run.dt / FS, "time_loop_s": t_loop}
    return PeriodicResult(r, t, mx, my, prov, warnings)


# --- emission --------------------------------------------------------------------------------------------------

@dataclass
class EmissionResult:
    wavelengths: np.ndarray
    p_tot: np.ndarray  # power the dipole gives to the field
    p_rad: np.ndarray  # power through the box around it
    provenance: dict = field(default_factory=dict)
    warnings: list = field(default_factory=list)


def solve_emission(device, layout: Layout, inv_eps, groups, media, wavelengths, pos, pdir, *, n_host: float = 1.0,
                   opts: DriverOptions | None = None) -> EmissionResult:
    """A point dipole at `pos` along `pdir` in the voxelised scene (emission.py:153-196): p_tot from the field at
    the source and the current's spectrum, p_rad from the box `layout.sca`."""
    opts = opts or DriverOptions()
    warnings: list[str] = []
    wl = np.asarray(wavelengths, dtype=float)
    p = plan_scattering(layout, media, wl, opts, n_host)
    dt, stride, dx = p["dt"], p["stride"], layout.dx
    omega = 2 * np.pi / wl
    src = PointDipole(shape=layout.shape, origin=layout.origin, dx=dx, dt=dt, pos=pos, pdir=pdir, inv_eps=inv_eps,
                      omega_c=p["omega_c"], omega_half=p["omega_half"], max_steps=p["max_steps"] + 4 * stride)
    box = flux_points(*layout.sca, layout.shape, dx)
    mon_src = DFTMonitor(omega, src.idx)
    mon_box = DFTMonitor(omega, box[0], box[1], box[2], energy=1)
    pml = PMLSpec(cells=((layout.pml, layout.pml),) * 3, kappa_max=opts.pml_kappa,
                  alpha_max=0.05 * n_host * float(omega.min()), eps_ref=n_host ** 2)
    eng = Engine(device, shape=layout.shape, dx=dx, dt=dt, single=True, inv_eps=inv_eps, groups=groups, pml=pml,
                 sources=[src], monitors=[mon_src, mon_box], dft_stride=stride)
    transit = float(np.max(layout.sca[1] - layout.sca[0])) * dx * n_host
    try:
        steps, level, graphs = run_until_decayed(eng, opts, max_steps=p["max_steps"],
                                                 min_steps=int((2 * src.t0 + transit) / dt),
                                                 window=math.ceil(2 * np.pi / omega.min() / (stride * dt)),
                                                 warnings=warnings)
        e_src, _ = mon_src.results()
        eb, hb = mon_box.results()
    finally:
        eng.close()
    j = src.spectrum(omega, stride, steps // stride)
    p_tot = -0.5 * dx ** 3 * np.real((e_src @ src.w) * np.conj(j))
    p_rad = flux(eb, hb, box[3])
    prov = {"steps": steps, "decay_level": level, "hip_graphs": graphs, "dft_stride": stride, "dt_fs": dt / FS}
    return EmissionResult(wl, p_tot, p_rad, prov, warnings)


__all__ = ["DriverOptions", "EmissionResult", "Layout", "PeriodicCell", "PeriodicResult", "ScatteringResult", "Sphere",
           "band", "e_node_positions", "flat_index", "periodic_time_step", "plan_scattering", "scattering_layout", "solve_emission",
           "solve_periodic", "solve_scattering", "staircase"]
