"""Matched configurations for the oracle (FP64 NumPy) and the HIP engine.

Each builder takes a namespace ``ns`` -- the oracle modules or the product package -- so that the very
same inputs (material arrays, index tables, source tables) reach both sides.
"""

from __future__ import annotations

import types

import numpy as np


def oracle_ns():
    from oracle import fdtd, media, sources
    return types.SimpleNamespace(Engine=fdtd.Engine, PMLSpec=fdtd.PMLSpec, DFTMonitor=fdtd.DFTMonitor,
                                 TimeMonitor=fdtd.TimeMonitor, DispersiveGroup=fdtd.DispersiveGroup,
                                 OffDiagonal=fdtd.OffDiagonal, PointDipole=sources.PointDipole,
                                 TFSFPlaneWave=sources.TFSFPlaneWave, TableSource=sources.TableSource,
                                 BlochPlaneSource=sources.BlochPlaneSource, ModeInjection=sources.ModeInjection,
                                 Medium=media.Medium, Pole=media.Pole, flat_index=fdtd.flat_index,
                                 e_region=fdtd.e_region, h_region=fdtd.h_region, TensorMesh=_oracle_tensor_mesh(),
                                 oracle=True)


def _oracle_tensor_mesh():
    from oracle.tensor import TensorMesh
    return TensorMesh


def product_ns():
    from lightsmith_amd import fdtd
    from lightsmith_amd.fdtd import sources
    return types.SimpleNamespace(Engine=fdtd.Engine, PMLSpec=fdtd.PMLSpec, DFTMonitor=fdtd.DFTMonitor,
                                 TimeMonitor=fdtd.TimeMonitor, DispersiveGroup=fdtd.DispersiveGroup,
                                 OffDiagonal=fdtd.OffDiagonal, PointDipole=sources.PointDipole,
                                 TFSFPlaneWave=sources.TFSFPlaneWave, TableSource=sources.TableSource,
                                 BlochPlaneSource=sources.BlochPlaneSource, ModeInjection=sources.ModeInjection,
                                 Medium=fdtd.Medium, Pole=fdtd.Pole, flat_index=fdtd.flat_index,
                                 e_region=fdtd.e_region, h_region=fdtd.h_region, TensorMesh=fdtd.TensorMesh,
                                 oracle=False)


def make(ns, device=None, **kw):
    """Engine of namespace ns; the oracle takes no device, the HIP engine runs float32 on `device`."""
    if ns.oracle:
        kw.pop("single", None)
        return ns.Engine(**kw)
    kw.setdefault("single", True)
    return ns.Engine(device, **kw)


def fields(eng):
    """(E, H) as complex128 NumPy arrays in the reference layout (3, N0, N1, N2)."""
    E, H = eng.E, eng.H
    if hasattr(E, "cpu"):
        E, H = E.cpu().numpy(), H.cpu().numpy()
    return np.asarray(E, dtype=complex), np.asarray(H, dtype=complex)


def rel_l2(a, b, mask=None) -> float:
    a, b = np.asarray(a), np.asarray(b)
    if mask is not None:
        a, b = a[mask], b[mask]
    den = np.linalg.norm(b.ravel())
    return float(np.linalg.norm((a - b).ravel()) / (den if den > 0 else 1.0))


def real_mask(shape, periodic):
    """(3, N0, N1, N2) mask of real (non-ghost) cells: ghost planes are not materialised by the HIP engine."""
    m = np.ones((3,) + tuple(shape), bool)
    for a, p in enumerate(periodic):
        if p:
            sl = [slice(None)] * 4
            sl[a + 1] = 0
            m[tuple(sl)] = False
            sl[a + 1] = shape[a] - 1
            m[tuple(sl)] = False
    return m


def random_fields(eng, seed: int, periodic=(False, False, False)):
    """Seed E and H with random values on their stepped regions (same values on both sides)."""
    rng = np.random.default_rng(seed)
    shape = eng.shape
    for c in range(3):
        for F, R in ((eng.E, eng.RE[c]), (eng.H, eng.RH[c])):
            sub = tuple(R)
            vals = rng.standard_normal(tuple(r.stop - r.start for r in sub))
            if hasattr(F, "cpu"):
                import torch
                F[c][sub] = torch.as_tensor(vals, dtype=F.dtype, device=F.device)
            else:
                F[c][sub] = vals
    return shape


# --- configurations -------------------------------------------------------------------------------

def lorentz3(rng, n_poles=3):
    """An n-pole (3 by default) Lorentz medium with seeded poles (omega_p, omega_0 in [1, 10] eV, gamma in
    [0.05, 0.5] eV)."""
    from oracle.media import HBARC_EV_NM
    poles = []
    for _ in range(n_poles):
        wp, w0, g = rng.uniform(1, 10), rng.uniform(1, 10), rng.uniform(0.05, 0.5)
        poles.append(((wp / HBARC_EV_NM) ** 2, w0 / HBARC_EV_NM, g / HBARC_EV_NM))
    return poles


def medium(ns, eps_inf, poles):
    return ns.Medium(eps_inf, tuple(ns.Pole(*p) for p in poles))


def dispersive_box(ns, device=None, shape=(40, 36, 32), dx=2.0, seed=0, pml_cells=6, n_steps_hint=400,
                   with_dipole=True, frac=0.1, kappa_max=1.0, alpha=0.0, monitors=True, material="auto", fused=None,
                   stride=10, pole_counts=None):
    """CPML box with random dielectric values, a fraction of D-form 3-pole Lorentz nodes, a z dipole and a
    DFT plane: a small version of the 320^3 throughput configuration. pole_counts (e.g. (1, 3)): one Lorentz
    medium of each count instead, each D-form cell given one of them at random (one pure group per medium)."""
    rng = np.random.default_rng(seed)
    eps = rng.choice([1.0, 2.25, 4.0, 12.0], size=shape)
    inv_eps = np.broadcast_to(1.0 / eps, (3,) + shape).copy()
    poles = lorentz3(np.random.default_rng(seed + 100))
    med = medium(ns, 1.0, poles)
    pml = tuple((pml_cells, pml_cells) for _ in range(3))
    inner = np.zeros(shape, bool)
    inner[pml_cells:-pml_cells, pml_cells:-pml_cells, pml_cells:-pml_cells] = True
    disp = inner & (rng.random(shape) < frac)
    ii, jj, kk = np.nonzero(disp)
    idx = np.concatenate([ns.flat_index(c, ii, jj, kk, shape) for c in range(3)])
    inv_eps.reshape(3, -1)[:, np.ravel_multi_index((ii, jj, kk), shape)] = 1.0
    groups = [ns.DispersiveGroup(idx=idx, media=[med], frac=np.ones((1, len(idx))), w=None)] if len(idx) else []
    if pole_counts:
        prng = np.random.default_rng(seed + 200)
        which = prng.integers(0, len(pole_counts), size=len(ii))
        groups = []
        for q, npl in enumerate(pole_counts):
            sel = which == q
            if not sel.any():
                continue
            m = medium(ns, 1.0, lorentz3(np.random.default_rng(seed + 300 + q), npl))
            gi = np.sort(np.concatenate([ns.flat_index(c, ii[sel], jj[sel], kk[sel], shape) for c in range(3)]))
            groups.append(ns.DispersiveGroup(idx=gi, media=[m], frac=np.ones((1, len(gi))), w=None))
    dt = 0.95 * dx / np.sqrt(3)
    sources = []
    if with_dipole:
        ctr = np.asarray(shape) * dx / 2 + 0.3
        sources.append(ns.PointDipole(shape=shape, origin=np.zeros(3), dx=dx, dt=dt, pos=ctr, pdir=(0.2, 0.1, 1.0),
                                      inv_eps=inv_eps, omega_c=2 * np.pi / 60, omega_half=0.15 * 2 * np.pi / 60,
                                      max_steps=n_steps_hint + 10))
    mons = []
    if monitors:
        kz = shape[2] // 2 + 3
        i = np.arange(pml_cells, shape[0] - pml_cells)
        j = np.arange(pml_cells, shape[1] - pml_cells)
        gi, gj = np.meshgrid(i, j, indexing="ij")
        gi, gj = gi.ravel(), gj.ravel()
        e = np.concatenate([ns.flat_index(0, gi, gj, kz, shape), ns.flat_index(1, gi, gj, kz, shape)])
        h1 = np.concatenate([ns.flat_index(1, gi, gj, kz - 1, shape), ns.flat_index(0, gi, gj, kz - 1, shape)])
        h2 = np.concatenate([ns.flat_index(1, gi, gj, kz, shape), ns.flat_index(0, gi, gj, kz, shape)])
        omega = 2 * np.pi / np.linspace(50, 80, 5)
        mons = [ns.DFTMonitor(omega, e, h1, h2, energy=1)]
    spec = ns.PMLSpec(cells=pml, kappa_max=kappa_max, alpha_max=alpha, eps_ref=1.0)
    kw = dict(shape=shape, dx=dx, dt=dt, inv_eps=inv_eps, groups=groups, pml=spec, sources=sources, monitors=mons,
              dft_stride=stride)
    if not ns.oracle:
        kw["material"] = material
        if fused is not None:
            kw["fused"] = fused
    return make(ns, device, **kw)


def tfsf_box(ns, device=None, shape=(30, 28, 34), khat=(0, 0, 1.0), ehat=(1.0, 0, 0), steps=300, particle=True):
    """A 1D-fed TFSF box around a small dielectric block, CPML outside."""
    dx = 1.0
    dt = 0.95 / np.sqrt(3)
    inv = np.full((3,) + shape, 1 / 1.77)
    if particle:
        inv[:, 13:18, 12:17, 15:20] = 1 / 6.0
    k = np.asarray(khat, float)
    k /= np.linalg.norm(k)
    src = ns.TFSFPlaneWave(shape=shape, origin=np.zeros(3), dx=dx, dt=dt, r0=np.array([15.0, 14, 17]),
                           lo=np.array([9, 8, 10]), hi=np.array([21, 20, 24]), khat=k, ehat=ehat, eps_b=1.77,
                           omega_c=0.6, omega_half=0.2, max_steps=steps + 5)
    mon = ns.DFTMonitor(np.array([0.5, 0.6]), *src.probe(), source=src)
    return make(ns, device, shape=shape, dx=dx, dt=dt, inv_eps=inv,
                pml=ns.PMLSpec(cells=((5, 5),) * 3, eps_ref=1.77), sources=[src], monitors=[mon], dft_stride=10)


def synthetic_table_source(ns, shape, seed=0, m=3, n_rows=40, n_pts=None):
    """A table-fed source over the TFSF faces of a box with random smooth tables (the per-step part only):
    every face point of each kind reads its own row at its own delay."""
    from oracle.sources import tfsf_slabs
    rng = np.random.default_rng(seed)
    L = 200
    n_c = L - 8
    t = np.arange(n_c)
    rows = np.stack([np.sin(0.07 * t * (1 + 0.1 * q)) * np.exp(-((t - 90) / 60.0) ** 2) for q in range(n_rows)])
    T = np.concatenate([rows, np.zeros((n_rows, 8))], axis=1)
    cmax = n_c + 2
    points = {}
    for kind in ("E", "H"):
        nodes, base, i0, u0, coef = [], [], [], [], []
        for k_, c, region, g, cf, q in tfsf_slabs(shape, (8, 8, 9), (20, 19, 23), 0.5, 1.0):
            if k_ != kind:
                continue
            ii, jj, kk = np.meshgrid(*[np.arange(r.start, r.stop) for r in region], indexing="ij")
            nn = ns.flat_index(c, ii, jj, kk, shape).reshape(-1)
            nodes.append(nn)
            r_ = rng.integers(0, n_rows, len(nn))
            base.append(r_ * L - 2)
            i0.append(rng.integers(6 * m, 12 * m, len(nn)))
            u0.append(rng.uniform(0, 1, len(nn)))
            coef.append(np.full(len(nn), cf) * rng.uniform(0.5, 1.5, len(nn)))
        points[kind] = tuple(np.concatenate(v) for v in (nodes, base, i0, u0, coef))
    inc = (np.arange(3) * L - 2, np.full(3, 6 * m + 1), np.full(3, 0.25))
    return ns.TableSource(T.reshape(-1), m, cmax, points, inc=inc, T_inc=T[:3].reshape(-1))


# --- adjoint ------------------------------------------------------------------------------------------

def adjoint_case(shape=(34, 28, 22), pml=6, steps=240, stride=8, seed=4, design=((13, 9, 8), (21, 19, 14)),
                 wavelength=400.0, monitor="plane", ade=False):
    """A small SOI-like adjoint problem: clad eps 1.44^2 with a core bar along x (eps 3.48^2), CPML with kappa
    and alpha, a mode-injection soft source on an x plane (E_y, Gaussian profile, Gaussian pulse), a design box
    of the given cells, and a two-frequency DFT monitor (E_y, E_z and face-averaged H_z, H_y) on an x plane
    downstream. Returns a dict of host inputs shared by the oracle and the HIP problem."""
    rng = np.random.default_rng(seed)
    n0, n1, n2 = shape
    dx = 20.0
    dt = 0.95 * dx / np.sqrt(3)
    e_clad, e_core = 1.44 ** 2, 3.48 ** 2
    eps = np.full(shape, e_clad)
    jc, kc = n1 // 2, n2 // 2
    eps[:, jc - 3:jc + 3, kc - 2:kc + 2] = e_core
    inv_eps = np.broadcast_to(1.0 / eps, (3,) + shape).copy()
    lo, hi = design
    # source: E_y on plane i = pml + 2 over a window around the core
    i_src = pml + 2
    jj, kk = np.meshgrid(np.arange(jc - 6, jc + 6), np.arange(kc - 5, kc + 5), indexing="ij")
    from oracle.fdtd import flat_index
    e_src = flat_index(1, np.full(jj.size, i_src), jj.ravel(), kk.ravel(), shape)
    prof = np.exp(-((jj.ravel() - jc + 0.5) / 4.0) ** 2 - ((kk.ravel() - kc + 0.5) / 3.0) ** 2)
    omega_c = 2 * np.pi / wavelength
    tau = 2.0 / (0.2 * omega_c)
    t = (np.arange(steps + 2) + 0.5) * dt
    pulse = np.exp(-((t - 3.5 * tau) / tau) ** 2) * np.sin(omega_c * (t - 3.5 * tau))
    # monitor on plane i = n0 - pml - 3
    i_mon = n0 - pml - 3
    jm, km = np.meshgrid(np.arange(jc - 7, jc + 7), np.arange(kc - 6, kc + 6), indexing="ij")
    jm, km = jm.ravel(), km.ravel()
    im = np.full(jm.size, i_mon)
    e_mon = np.concatenate([flat_index(1, im, jm, km, shape), flat_index(2, im, jm, km, shape)])
    h1 = np.concatenate([flat_index(2, im - 1, jm, km, shape), flat_index(1, im - 1, jm, km, shape)])
    h2 = np.concatenate([flat_index(2, im, jm, km, shape), flat_index(1, im, jm, km, shape)])
    if monitor == "box":
        # a closed flux box downstream of the design: its edge and corner nodes are shared by two or three faces
        from lightsmith_amd.fdtd.run import flux_points
        e_mon, h1, h2, _ = flux_points((i_mon - 3, jc - 5, kc - 4), (i_mon + 1, jc + 5, kc + 4), shape, dx)
    omega = omega_c * np.array([0.95, 1.05])
    rho = rng.uniform(0, 1, size=tuple(h - l for l, h in zip(lo, hi)))
    ade_idx, ade_medium = None, None
    if ade:
        # a Drude + Lorentz block (silver-like: eps_inf 2, wp 9 eV, one Lorentz pole at 4 eV) in the core between
        # the design box and the monitor plane: pure D-form nodes (inv_eps 1) whose update the adjoint transposes
        from oracle.media import HBARC_EV_NM
        w0 = 4.0 / HBARC_EV_NM
        ade_medium = (2.0, ((9.0 / HBARC_EV_NM) ** 2, 0.0, 0.05 / HBARC_EV_NM), (1.5 * w0 ** 2, w0, 0.5 / HBARC_EV_NM))
        blk = (slice(hi[0] + 1, hi[0] + 3), slice(jc - 3, jc + 3), slice(kc - 2, kc + 2))
        ii, jj2, kk2 = np.meshgrid(*[np.arange(b.start, b.stop) for b in blk], indexing="ij")
        ade_idx = np.sort(np.concatenate([flat_index(c, ii.ravel(), jj2.ravel(), kk2.ravel(), shape) for c in range(3)]))
        inv_eps.reshape(-1)[ade_idx] = 1.0
    w_e = rng.standard_normal(len(e_mon)) + 1j * rng.standard_normal(len(e_mon))
    w_h = rng.standard_normal(len(h1)) + 1j * rng.standard_normal(len(h1))
    return dict(shape=shape, dx=dx, dt=dt, inv_eps=inv_eps, pml=pml, steps=steps, stride=stride, lo=lo, hi=hi,
                eps_lo=e_clad, eps_hi=e_core, e_src=e_src, m_src=prof[:, None], t_src=pulse[:, None], omega=omega,
                e_mon=e_mon, h1=h1, h2=h2, rho=rho, w_e=w_e, w_h=w_h, ade_idx=ade_idx, ade_medium=ade_medium)


def adjoint_objective(acc_e, acc_h, case):
    """|sum w_e acc_e[0]|^2 + |sum w_h acc_h[1]|^2 (numpy or torch)."""
    if isinstance(acc_e, np.ndarray):
        return abs(np.sum(case["w_e"] * acc_e[0])) ** 2 + abs(np.sum(case["w_h"] * acc_h[1])) ** 2
    import torch
    we = torch.as_tensor(case["w_e"], device=acc_e.device).to(acc_e.dtype)
    wh = torch.as_tensor(case["w_h"], device=acc_h.device).to(acc_h.dtype)
    return (we * acc_e[0]).sum().abs() ** 2 + (wh * acc_h[1]).sum().abs() ** 2


def adjoint_inv_eps(case, rho):
    """inv_eps of the case with the design box set from rho (the staircase rule of DesignRegion)."""
    inv = case["inv_eps"].copy()
    box = tuple(slice(l, h) for l, h in zip(case["lo"], case["hi"]))
    eps = case["eps_lo"] + np.asarray(rho) * (case["eps_hi"] - case["eps_lo"])
    for c in range(3):
        inv[c][box] = 1.0 / eps
    return inv


def adjoint_oracle(case, rho):
    """The FP64 oracle Engine of the case at density rho (not stepped)."""
    from oracle import fdtd, sources
    p = case["pml"]
    src = sources.ModeInjection(np.zeros(0, np.int64), np.zeros((0, 1)), np.zeros((case["steps"] + 2, 1)), case["e_src"],
                                case["m_src"], case["t_src"])
    mon = fdtd.DFTMonitor(case["omega"], case["e_mon"], case["h1"], case["h2"])
    return fdtd.Engine(shape=case["shape"], dx=case["dx"], dt=case["dt"], inv_eps=adjoint_inv_eps(case, rho),
                       pml=fdtd.PMLSpec(cells=((p, p),) * 3, kappa_max=case.get("kappa_max", 2.0),
                                        alpha_max=case.get("alpha_max", 0.02), eps_ref=case["eps_lo"]),
                       sources=[src], monitors=[mon], dft_stride=case["stride"],
                       groups=_adjoint_groups(oracle_ns(), case))


def _adjoint_groups(ns, case):
    """The case's pure D-form group (none without ade) in namespace ns."""
    if case.get("ade_idx") is None:
        return []
    eps_inf, *poles = case["ade_medium"]
    med = ns.Medium(eps_inf, tuple(ns.Pole(*p) for p in poles))
    idx = case["ade_idx"]
    return [ns.DispersiveGroup(idx=idx, media=[med], frac=np.ones((1, len(idx))), w=None)]


def adjoint_problem(case, device, **kw):
    """The HIP AdjointProblem of the case."""
    from lightsmith_amd.fdtd import DFTMonitor, PMLSpec
    from lightsmith_amd.fdtd.adjoint import AdjointProblem, DesignRegion
    from lightsmith_amd.fdtd.sources import ModeInjection
    p = case["pml"]
    src = ModeInjection(np.zeros(0, np.int64), np.zeros((0, 1)), np.zeros((case["steps"] + 2, 1)), case["e_src"],
                        case["m_src"], case["t_src"])
    mon = DFTMonitor(case["omega"], case["e_mon"], case["h1"], case["h2"])
    return AdjointProblem(device, shape=case["shape"], dx=case["dx"], dt=case["dt"], inv_eps=case["inv_eps"],
                          pml=PMLSpec(cells=((p, p),) * 3, kappa_max=2.0, alpha_max=0.02, eps_ref=case["eps_lo"]),
                          sources=[src], monitor=mon,
                          design=DesignRegion(case["lo"], case["hi"], case["eps_lo"], case["eps_hi"]),
                          n_steps=case["steps"], dft_stride=case["stride"], groups=_adjoint_groups(product_ns(), case),
                          **kw)


# --- conformal vertex tensor ---------------------------------------------------------------------------

def tensor_sphere(ns, device=None, shape=(24, 24, 24), steps=300, periodic=(False, False, False), bloch=(0.0, 0.0, 0.0),
                  pml_cells=4, center=(0.3, 0.1, -0.2), radius=6.4):
    """A Drude-metal sphere (eps_inf 9.5, omega_p = 9 eV) in water (eps 1.77), meshed with the conformal vertex
    tensor (the oracle's restatement of setup.vertex_tensor) around the interface: staircase water / D-form metal
    nodes elsewhere, tensor nodes at inv_eps = 1. A z dipole next to the sphere, CPML on the non-periodic axes,
    a DFT monitor on an x plane."""
    from oracle import media as omed
    from oracle.tensor import vertex_tensor
    o_media = [omed.Medium(1.77, ()), omed.Medium(9.5, (omed.Pole(0.00208, 0.0, 1.8e-4),))]
    media = [ns.Medium(1.77, ()), ns.Medium(9.5, (ns.Pole(0.00208, 0.0, 1.8e-4),))]
    c = np.asarray(center)
    dx = 1.0
    origin = -np.asarray(shape, float) / 2 * dx
    slot_fn = lambda p: (np.linalg.norm(p - c, axis=-1) < radius).astype(int)  # noqa: E731
    roi_lo, roi_hi = np.full(3, 2), np.asarray(shape) - 2
    tm = vertex_tensor(slot_fn, o_media, shape, origin, dx, roi_lo, roi_hi, periodic=periodic)
    assert tm is not None and len(tm.normal) > 50
    n0, n1, n2 = shape
    inv_eps = np.empty((3,) + shape)
    metal = []
    for comp in range(3):
        off = np.zeros(3)
        off[comp] = 0.5
        ii, jj, kk = np.meshgrid(*[np.arange(n) for n in shape], indexing="ij")
        pos = origin + (np.stack([ii, jj, kk], -1) + off) * dx
        m = slot_fn(pos.reshape(-1, 3)).reshape(shape)
        inv_eps[comp] = np.where(m == 1, 1.0, 1 / 1.77)
        mi = np.nonzero(m == 1)
        metal.append(ns.flat_index(comp, *mi, shape))
    metal = np.concatenate(metal)
    taken = tm.idx
    inv_eps.reshape(-1)[taken] = 1.0
    metal = metal[~np.isin(metal, taken)]
    tm_ns = ns.TensorMesh(media=media, frac=tm.frac, normal=tm.normal, node_comp=tm.node_comp,
                          node_idx3=tm.node_idx3, vert_nodes=tm.vert_nodes, shift=tm.shift, slot_vert=tm.slot_vert,
                          slot_medium=tm.slot_medium, shape=tm.shape, vert_idx3=tm.vert_idx3)
    groups = [ns.DispersiveGroup(idx=metal, media=[media[1]], frac=np.ones((1, len(metal))), w=None), tm_ns]
    dt = 0.5
    pos = c + np.array([radius + 2.3, 0.4, 0.2])
    src = ns.PointDipole(shape=shape, origin=origin, dx=dx, dt=dt, pos=pos, pdir=(0.1, 0.2, 1.0), inv_eps=inv_eps,
                         omega_c=2 * np.pi / 25.0, omega_half=0.3 * 2 * np.pi / 25.0, max_steps=steps + 10)
    cells = tuple((0, 0) if p else (pml_cells, pml_cells) for p in periodic)
    kx = n0 // 2 + 5
    jj, kk = np.meshgrid(np.arange(3, n1 - 3), np.arange(3, n2 - 3), indexing="ij")
    e_idx = np.concatenate([ns.flat_index(cc, np.full(jj.size, kx), jj.ravel(), kk.ravel(), shape) for cc in (1, 2)])
    mon = ns.DFTMonitor(2 * np.pi / np.array([22.0, 28.0]), e_idx)
    kw = dict(shape=shape, dx=dx, dt=dt, inv_eps=inv_eps, groups=groups,
              pml=ns.PMLSpec(cells=cells, eps_ref=1.77), sources=[src], monitors=[mon], dft_stride=5,
              periodic=periodic, bloch=bloch)
    return make(ns, device, **kw)


def oracle_scattering_engine(scene, dtype=np.float64):
    """The FP64 oracle Engine of a lightsmith_amd.fdtd.scattering.Scene: the same material arrays, groups (the
    vertex tensor as the oracle's TensorMesh), off-diagonal rows, CPML, source tables and monitor index sets."""
    from oracle import fdtd, sources
    from oracle.tensor import TensorMesh
    groups = []
    for g in scene.groups:
        if hasattr(g, "slot_vert"):
            g = TensorMesh(media=g.media, frac=g.frac, normal=g.normal, node_comp=g.node_comp, node_idx3=g.node_idx3,
                           vert_nodes=g.vert_nodes, shift=g.shift, slot_vert=g.slot_vert, slot_medium=g.slot_medium,
                           shape=g.shape, vert_idx3=g.vert_idx3)
        groups.append(g)
    s = scene.source
    src = sources.TableSource(s.T.reshape(-1), s.m, s.cmax, s.points, inc=s.inc, T_inc=np.asarray(s.T_inc).reshape(-1))
    p = scene.pml
    pml = fdtd.PMLSpec(cells=p.cells, m=p.m, kappa_max=p.kappa_max, alpha_max=p.alpha_max, eps_ref=p.eps_ref)
    mons = [fdtd.DFTMonitor(scene.omega, scene.sca[0], scene.sca[1], scene.sca[2], energy=1),
            fdtd.DFTMonitor(scene.omega, scene.tot[0], scene.tot[1], scene.tot[2], energy=2),
            fdtd.DFTMonitor(scene.omega, np.arange(3), source=src)]
    lay = scene.layout
    od = scene.offdiag
    return fdtd.Engine(shape=lay.shape, dx=lay.dx, dt=scene.dt, inv_eps=scene.inv_eps, groups=groups, pml=pml,
                       sources=[src], monitors=mons, dft_stride=scene.stride, dtype=dtype,
                       offdiag=None if od is None else fdtd.OffDiagonal(od.idx3, od.comp, od.coef))
