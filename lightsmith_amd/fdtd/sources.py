"""Sources of the HIP engine, following the reference's source classes.

Host set-up (geometry, coefficients, time tables) runs in float64 NumPy at construction; ``bind(eng)``
uploads float32 tables and registers the per-step work with the native engine, which then performs
what the reference's ``correct_H``/``advance_H``/``correct_E``/``advance_E`` hooks do, in the same
order within the step (engine.py:665-678). The Python hooks are therefore no-ops here.
"""

from __future__ import annotations

import numpy as np
import torch

from .._lib import PH_ADVANCE_E, PH_CORRECT_E, PH_CORRECT_H, check, ptr
from .engine import Engine, e_region, flat_index, h_region

PML1D = 48  # absorbing cells at each end of the 1D auxiliary grid (tfsf.py:22)


class _Native:
    """Hook protocol stubs: the per-step work runs inside the engine's native step."""

    native = True

    def correct_H(self, eng):
        pass

    def advance_H(self, eng):
        pass

    def correct_E(self, eng):
        pass

    def advance_E(self, eng):
        pass

    def bind_native(self, eng):  # pragma: no cover - overridden
        raise NotImplementedError

    def bind(self, eng):
        self.bind_native(eng)


def gaussian_sine(t, t0: float, tau: float, omega_c: float) -> np.ndarray:
    """exp(-((t - t0) / tau)^2) sin(omega_c (t - t0))."""
    return np.exp(-(((t - t0) / tau) ** 2)) * np.sin(omega_c * (t - t0))


# --- point dipole (emission.py:97-150) -----------------------------------------------------------

class PointDipole(_Native):
    """A point current J = j(t) p_hat at ``pos`` spread trilinearly over the eight E_c nodes around it;
    E_i -= dt (1/eps)_i w_i j(t) after the E curl and before the D-form groups. j is the exact difference
    of a Gaussian-modulated sine over omega_c dt, so it carries no net charge (emission.py:120-127)."""

    def __init__(self, *, shape, origin, dx: float, dt: float, pos, pdir, inv_eps, omega_c: float,
                 omega_half: float, max_steps: int):
        origin = np.asarray(origin, dtype=float)
        pos = np.asarray(pos, dtype=float)
        idx, w = [], []
        for c in range(3):
            if abs(pdir[c]) < 1e-12:
                continue
            off = np.zeros(3)
            off[c] = 0.5
            f = (pos - origin) / dx - off
            base = np.floor(f).astype(int)
            frac = f - base
            for corner in range(8):
                bits = np.array([(corner >> a) & 1 for a in range(3)])
                wt = float(np.prod(np.where(bits, frac, 1 - frac)))
                if wt > 1e-12:
                    idx.append(int(flat_index(c, *(base + bits), shape)))
                    w.append(wt * pdir[c])
        self.idx, self.w = np.asarray(idx, dtype=np.int64), np.asarray(w, dtype=float)
        self.coef = dt * np.asarray(inv_eps, dtype=float).reshape(-1)[self.idx] * self.w
        self.tau = 2.0 / omega_half
        self.t0 = 5.0 * self.tau
        g = gaussian_sine(np.arange(max_steps + 3) * dt, self.t0, self.tau, omega_c)
        self.wave = np.diff(g) / (omega_c * dt)
        self.dt = dt

    def bind_native(self, eng: Engine):
        eng.add_modal(PH_CORRECT_E, "E", eng.native_index(self.idx), -self.coef, self.wave)

    def spectrum(self, omega, stride: int, n_samples: int) -> np.ndarray:
        """j(w) sampled as the DFT monitors sample E (emission.py:142-150)."""
        N = stride * np.arange(1, n_samples + 1)
        t = (N - 0.5) * self.dt
        return np.exp(1j * np.outer(omega, t)) @ self.wave[N - 1]


# --- TFSF box geometry (tfsf.py:25-105) ----------------------------------------------------------

def _offsets(kind: str, c: int) -> np.ndarray:
    o = np.full(3, 0.5) if kind == "H" else np.zeros(3)
    o[c] = 0.0 if kind == "H" else 0.5
    return o


def matched_eps_1d(khat, omega: float, dt: float, dx: float, eps_b: float) -> float:
    """Permittivity of a 1D Yee line whose phase velocity matches the 3D grid's along khat at omega
    (tfsf.py:35-47): bisect the 3D dispersion relation for |k|, then invert the 1D one."""
    khat = np.asarray(khat, dtype=float)
    target = eps_b * (np.sin(omega * dt / 2) / dt) ** 2
    lo, hi = 0.0, np.pi / dx
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if np.sum((np.sin(mid * khat * dx / 2) / dx) ** 2) < target:
            lo = mid
        else:
            hi = mid
    k = 0.5 * (lo + hi)
    return float((np.sin(k * dx / 2) / dx) ** 2 / (np.sin(omega * dt / 2) / dt) ** 2)


def tfsf_slabs(shape, lo, hi, dt: float, spacing, periodic=(False, False, False)):
    """One-cell correction slabs of a TFSF box [lo, hi] (node units), tfsf.py:50-105. Yields
    (kind, c, region, g, coef, q): kind_c in region was updated from a neighbour across a box face and
    gets += coef * F_g(q), F the incident H (kind "E") or E (kind "H"), q the neighbour's index position."""
    spacing = np.broadcast_to(np.asarray(spacing, dtype=float), (3,))
    lo, hi = np.asarray(lo, dtype=float), np.asarray(hi, dtype=float)
    out = []
    for kind in ("E", "H"):
        for c in range(3):
            R = e_region(c, shape, periodic) if kind == "E" else h_region(c, shape, periodic)
            off = _offsets(kind, c)
            d1, d2 = (c + 1) % 3, (c + 2) % 3
            for d, g, s in ((d1, d2, 1.0), (d2, d1, -1.0)):
                s = -s if kind == "H" else s
                across = {}
                for a in range(3):
                    if a != d:
                        ix = np.arange(R[a].start, R[a].stop)
                        across[a] = ix[(ix + off[a] >= lo[a]) & (ix + off[a] <= hi[a])]
                if any(len(v) == 0 for v in across.values()):
                    continue
                along = np.arange(R[d].start, R[d].stop)
                u = along + off[d]
                inside = ((u >= lo[d]) & (u <= hi[d])).astype(int)
                for sign_q, du in ((1.0, 0.5), (-1.0, -0.5)):
                    neigh = ((u + du >= lo[d]) & (u + du <= hi[d])).astype(int)
                    delta = inside - neigh
                    for i in np.nonzero(delta)[0]:
                        region, axes = [None] * 3, []
                        for a in range(3):
                            if a == d:
                                region[a] = slice(int(along[i]), int(along[i]) + 1)
                                axes.append(np.array([u[i] + du]))
                            else:
                                region[a] = slice(int(across[a][0]), int(across[a][-1]) + 1)
                                axes.append(across[a] + off[a])
                        q = np.stack(np.meshgrid(*axes, indexing="ij"), axis=-1)
                        out.append((kind, c, tuple(region), g, s * sign_q * delta[i] * (dt / spacing[d]), q))
    return out


def region_nodes(shape, c: int, region) -> np.ndarray:
    """Reference flat indices of component c over an index box."""
    ii, jj, kk = np.meshgrid(*[np.arange(r.start, r.stop) for r in region], indexing="ij")
    return flat_index(c, ii, jj, kk, shape).reshape(-1)


# --- 1D-fed TFSF plane wave (tfsf.py:108-211) -----------------------------------------------------

class TFSFPlaneWave(_Native):
    """TFSF box fed by an auxiliary 1D Yee line along k (same dt and spacing), with graded absorbing ends
    and a Gaussian-modulated sine injected at m_src. Oblique incidence interpolates the line linearly,
    its permittivity matched to the 3D grid's phase velocity along k."""

    def __init__(self, *, shape, origin, dx, dt, r0, lo, hi, khat, ehat, eps_b, omega_c, omega_half,
                 max_steps: int, periodic=(False, False, False), m0: int | None = None):
        spacing = np.broadcast_to(np.asarray(dx, dtype=float), (3,)).copy()
        self.shape, self.dx, self.dt = tuple(shape), float(spacing[0]), float(dt)
        self.khat, self.ehat = np.asarray(khat, float), np.asarray(ehat, float)
        self.hhat = np.cross(self.khat, self.ehat)
        self.eps_b = float(eps_b)
        self.axis_aligned = int(np.sum(np.abs(self.khat) > 1e-12)) == 1
        if not self.axis_aligned and len(set(spacing)) > 1:
            raise ValueError("oblique TFSF incidence needs cubic cells")
        self.eps_1d = self.eps_b if self.axis_aligned else matched_eps_1d(self.khat, omega_c, dt, spacing[0], eps_b)
        origin, r0 = np.asarray(origin, float), np.asarray(r0, float)
        raw = []
        for kind, c, region, g, coef, q in tfsf_slabs(shape, lo, hi, dt, spacing, periodic):
            proj = self.hhat if kind == "E" else self.ehat
            if abs(proj[g]) < 1e-12:
                continue
            sq = (origin + q * spacing - r0) @ self.khat
            raw.append((kind, c, region, coef * proj[g] / (eps_b if kind == "E" else 1.0), sq))
        s_all = np.concatenate([r[4].ravel() for r in raw] + [np.zeros(1)])
        ds = float(spacing[int(np.argmax(np.abs(self.khat)))]) if self.axis_aligned else float(spacing[0])
        self.m_src = PML1D + 2
        self.m0 = max(int(self.m_src + 4 + np.ceil(-s_all.min() / ds)), m0 or 0)
        self.M = self.m0 + int(np.ceil(s_all.max() / ds)) + 4 + PML1D + 2
        self.slabs = []
        for kind, c, region, coef, sq in raw:
            f = sq.ravel() / ds + self.m0 - (0.5 if kind == "E" else 0.0)  # E corrections sample H1d
            i0 = np.floor(f + 1e-9).astype(np.int64)
            w = f - i0
            w[np.abs(w) < 1e-9] = 0.0
            self.slabs.append((kind, c, region, float(coef), i0, w if np.any(w) else None,
                               tuple(r.stop - r.start for r in region)))
        self.s_far = float(s_all.max())
        sig_max = (3 + 1) * np.log(1e9) / (2 * np.sqrt(self.eps_1d) * PML1D * ds)

        def sig(x):
            rho = np.clip(np.maximum((PML1D - x) / PML1D, (x - (self.M - 1 - PML1D)) / PML1D), 0, 1)
            return sig_max * rho ** 3

        se = sig(np.arange(self.M, dtype=float)) * dt / 2
        sh = sig(np.arange(self.M, dtype=float) + 0.5) * dt / 2
        self.ca, self.cb = (1 - se) / (1 + se), (dt / (self.eps_1d * ds)) / (1 + se)
        self.da, self.db = (1 - sh) / (1 + sh), (dt / ds) / (1 + sh)
        self.tau = 2.0 / omega_half
        self.t0 = 3.5 * self.tau
        self.wave = gaussian_sine((np.arange(max_steps + 2) + 1) * dt, self.t0, self.tau, omega_c)
        v = 1 / np.sqrt(self.eps_1d)
        self.t_passed = 2 * self.t0 + (self.s_far + (self.m0 - self.m_src) * ds) / v

    def probe(self):
        """Line indices (e, h1, h2) of the incident wave at s = 0 (tfsf.py:195-197)."""
        return np.array([self.m0]), np.array([self.m0 - 1]), np.array([self.m0])

    def bind_native(self, eng: Engine):
        if eng.complex:
            raise ValueError("the 1D-fed TFSF source drives real fields (use BlochPlaneSource under a Bloch phase)")
        dev = eng._dev
        self.E_flat = torch.zeros(self.M, dtype=torch.float32, device=dev)
        self.H_flat = torch.zeros(self.M, dtype=torch.float32, device=dev)
        f32 = lambda a: eng.to_device(np.asarray(a, np.float64))  # noqa: E731
        ca, cb, da, db = f32(self.ca), f32(self.cb), f32(self.da), f32(self.db)
        wave = f32(self.wave)
        eng._keep += [self.E_flat, self.H_flat]
        eng._state += [self.E_flat, self.H_flat]
        for kind, phase, src in (("H", PH_CORRECT_H, self.E_flat), ("E", PH_CORRECT_E, self.H_flat)):
            tg, si, ww = [], [], []
            for k, c, region, coef, i0, w, shape_ in self.slabs:
                if k != kind:
                    continue
                nodes = eng.native_index(region_nodes(self.shape, c, region))
                wl = np.zeros(len(i0)) if w is None else w
                tg.append(nodes)
                si.append(np.stack([i0, i0 + 1], axis=1))
                ww.append(np.stack([coef * (1 - wl), coef * wl], axis=1))
            if tg:
                eng.add_gather(phase, kind, np.concatenate(tg), src, np.concatenate(si), np.concatenate(ww))
        check(eng.lib.lsf_add_line1d(eng._h, self.M, ptr(self.E_flat), ptr(self.H_flat), ptr(ca), ptr(cb), ptr(da),
                                     ptr(db), self.m_src, ptr(wave), len(self.wave)), "lsf_add_line1d")
        self._eng_dev = dev

    def monitor_arrays(self):
        return self.E_flat, self.H_flat, 0, lambda idx: np.asarray(idx, dtype=np.int64)


# --- table-fed TFSF box, per step (layered.py:164-242, 472-482) -------------------------------------

STENCIL = np.arange(-2, 4)


class TableSource(_Native):
    """Per-step part of a table-fed TFSF box (ObliqueLayeredPlaneWave): every correction point reads its own
    table row at its own delay with quintic interpolation (tables sampled every m steps, read index clamped
    to cmax, zero rows past the end) and adds coef * value to its node. ``points[kind]`` = (reference flat
    node index, base, i0, u0, coef); ``inc`` = (base, i0, u0) of the incident-wave view and ``T_inc`` its
    tables (layered.py:469-482). The view's buffer (``E_flat``) is what monitors sample; it is evaluated when
    they sample (each chunk's end, ``Engine.sample``), not on every step."""

    def __init__(self, T, m: int, cmax: int, points: dict, inc=None, T_inc=None):
        self.T, self.m, self.cmax = np.asarray(T, dtype=float), int(m), int(cmax)
        self.points = points
        self.inc, self.T_inc = inc, T_inc

    def bind_native(self, eng: Engine):
        if eng.complex:
            raise ValueError("table-fed TFSF drives real fields")
        self._T = eng.to_device(self.T.reshape(-1).astype(np.float64))
        for kind, phase in (("H", PH_CORRECT_H), ("E", PH_CORRECT_E)):
            if kind not in self.points:
                continue
            nodes, base, i0, u0, coef = self.points[kind]
            eng.add_table(phase, kind, eng.native_index(nodes), self._T, self.m, self.cmax, base, i0, u0, coef)
        self.E_flat = self.H_flat = None
        if self.inc is not None:
            self.E_flat = eng.zeros(3, complex_=False)
            self.H_flat = self.E_flat
            self._Tinc = eng.to_device(np.asarray(self.T_inc, dtype=np.float64).reshape(-1))
            base, i0, u0 = (np.asarray(v) for v in self.inc)
            b, i, u = (eng.to_device(base.astype(np.int64)), eng.to_device(i0.astype(np.int64)),
                       eng.to_device(u0.astype(np.float32)))
            check(eng.lib.lsf_add_table_view(eng._h, PH_ADVANCE_E, len(base), ptr(self.E_flat), ptr(self._Tinc),
                                             self.m, self.cmax, ptr(b), ptr(i), ptr(u)), "lsf_add_table_view")

    def monitor_arrays(self):
        return self.E_flat, self.H_flat, 0, lambda idx: np.asarray(idx, dtype=np.int64)


# --- Bloch plane source (tfsf.py:214-296) -------------------------------------------------------------

class BlochPlaneSource(_Native):
    """One-sided TFSF plane (normal +z) for a Bloch-periodic cell at fixed (kx, ky), fed by an auxiliary
    column engine (one real cell in x and y, same Bloch phases, z spacing, CPML and dt) driven by a
    current sheet along s = (-Ky, Kx) or p = (Kx, Ky), K the grid wavenumbers. Total field for z-nodes
    >= k_src."""

    def __init__(self, *, shape, spacing, dt, k_src, k_launch, kpar, pol: str, eps_b, pml, omega_c,
                 omega_half, max_steps: int):
        self.shape, self.spacing, self.dt = tuple(shape), tuple(float(d) for d in spacing), float(dt)
        self.k_src, self.k_launch = int(k_src), int(k_launch)
        self.kpar = (float(kpar[0]), float(kpar[1]))
        self.eps_b, self.pml = float(eps_b), pml
        dx, dy = self.spacing[0], self.spacing[1]
        K = np.array([2 * np.sin(self.kpar[a] * self.spacing[a] / 2) / self.spacing[a] for a in (0, 1)])
        if not np.any(K):
            raise ValueError("the Bloch plane source needs an oblique k_par (use TFSFPlaneWave at normal incidence)")
        u = K / np.linalg.norm(K)
        a = np.array([-u[1], u[0]]) if pol == "s" else u
        self.amp = (a[0] * np.exp(0.5j * self.kpar[0] * dx), a[1] * np.exp(0.5j * self.kpar[1] * dy))
        nx, ny = shape[0] - 2, shape[1] - 2
        i, j = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
        self.phase = np.exp(1j * (self.kpar[0] * i * dx + self.kpar[1] * j * dy))
        self.tau = 2.0 / omega_half
        self.t0 = 3.5 * self.tau
        self.wave = gaussian_sine((np.arange(max_steps + 2) + 1) * dt, self.t0, self.tau, omega_c)
        self.t_passed = 2 * self.t0

    def probe(self, k: int):
        """Column flat indices (e, h1, h2) for S_z = Ex Hy - Ey Hx on node plane k (tfsf.py:248-254)."""
        n = (3, 3, self.shape[2])
        e = np.array([flat_index(0, 1, 1, k, n), flat_index(1, 1, 1, k, n)])
        h1 = np.array([flat_index(1, 1, 1, k - 1, n), flat_index(0, 1, 1, k - 1, n)])
        h2 = np.array([flat_index(1, 1, 1, k, n), flat_index(0, 1, 1, k, n)])
        return e, h1, h2

    def bind_native(self, eng: Engine):
        if not eng.complex:
            raise ValueError("the Bloch plane source needs complex (Bloch) fields")
        nz = self.shape[2]
        col = Engine(eng.device, shape=(3, 3, nz), dx=self.spacing, dt=self.dt, single=True,
                     inv_eps=np.full((3, 3, 3, nz), 1 / self.eps_b), groups=[], pml=self.pml, sources=[],
                     monitors=[], dft_stride=1, periodic=(True, True, False), bloch=(*self.kpar, 0.0), _sub=True,
                     _counter=eng._counter)
        self.col = col
        eng.add_subengine(col)
        nx, ny = self.shape[0] - 2, self.shape[1] - 2
        k = self.k_src
        cE = self.dt / (self.eps_b * self.spacing[2])
        cH = self.dt / self.spacing[2]
        ii, jj = np.meshgrid(np.arange(1, nx + 1), np.arange(1, ny + 1), indexing="ij")
        ph = self.phase.reshape(-1)
        plane = lambda c, kk: eng.native_index(flat_index(c, ii, jj, kk, self.shape).reshape(-1))  # noqa: E731
        colidx = lambda c, kk: col.native_index(np.array([flat_index(c, 1, 1, kk, (3, 3, nz))]))[0]  # noqa: E731
        npts = nx * ny
        # correct_H: H_x(k-1) -= cH phase E_y^col(k); H_y(k-1) += cH phase E_x^col(k)  (tfsf.py:274-279)
        tg = np.concatenate([plane(0, k - 1), plane(1, k - 1)])
        si = np.concatenate([np.full(npts, colidx(1, k)), np.full(npts, colidx(0, k))])
        ww = np.concatenate([-cH * ph, cH * ph])
        eng.add_gather(PH_CORRECT_H, "H", tg, col.E_flat, si[:, None], ww[:, None])
        # correct_E: E_x(k) += cE phase H_y^col(k-1); E_y(k) -= cE phase H_x^col(k-1)  (tfsf.py:281-286)
        tg = np.concatenate([plane(0, k), plane(1, k)])
        si = np.concatenate([np.full(npts, colidx(1, k - 1)), np.full(npts, colidx(0, k - 1))])
        ww = np.concatenate([cE * ph, -cE * ph])
        eng.add_gather(PH_CORRECT_E, "E", tg, col.H_flat, si[:, None], ww[:, None])
        # current sheet on the column after its E update (tfsf.py:291-296)
        kl = self.k_launch
        col.add_modal(PH_ADVANCE_E, "E", [colidx(0, kl), colidx(1, kl)],
                      np.array([[self.amp[0]], [self.amp[1]]]), self.wave)

    def monitor_arrays(self):
        return self.col.E_flat, self.col.H_flat, 1, self.col.native_index


# --- mode source injection (ports.py:362-372) -------------------------------------------------------

class ModeInjection(_Native):
    """H_flat[h] += sum_r mH[:, r] tE[n, r]; E_flat[e] += sum_r mE[:, r] tH[n, r] (ModeSource per step:
    rank-r real basis times time tables)."""

    def __init__(self, h_idx, mH, tE, e_idx, mE, tH):
        self.h_idx, self.mH, self.tE = np.asarray(h_idx), np.asarray(mH, float), np.asarray(tE, float)
        self.e_idx, self.mE, self.tH = np.asarray(e_idx), np.asarray(mE, float), np.asarray(tH, float)

    def bind_native(self, eng: Engine):
        if len(self.h_idx):
            eng.add_modal(PH_CORRECT_H, "H", eng.native_index(self.h_idx), self.mH, self.tE)
        if len(self.e_idx):
            eng.add_modal(PH_CORRECT_E, "E", eng.native_index(self.e_idx), self.mE, self.tH)
