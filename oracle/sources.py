"""Sources restated in FP64 (oracle; test infrastructure only).

Each class follows the reference's source hook protocol (engine.py:12-13, 665-678): ``bind(eng)``,
then per step ``correct_H``, ``advance_H``, ``correct_E``, ``advance_E``, reading the step counter
``eng.n`` (steps completed before the current one).
"""

from __future__ import annotations

import numpy as np

from .fdtd import Engine, e_region, flat_index, h_region

PML1D = 48  # tfsf.py:22


# --- point dipole (emission.py:97-150) ----------------------------------------------------------

class PointDipole:
    def __init__(self, *, shape, origin, dx, dt, pos, pdir, inv_eps, omega_c, omega_half, max_steps):
        idx, w = [], []
        origin = np.asarray(origin, dtype=float)
        for c in range(3):
            if abs(pdir[c]) < 1e-12:
                continue
            off = np.zeros(3)
            off[c] = 0.5
            f = (np.asarray(pos, dtype=float) - origin) / dx - off
            i0 = np.floor(f).astype(int)
            t = f - i0
            for corner in range(8):
                bits = np.array([(corner >> a) & 1 for a in range(3)])
                wt = float(np.prod(np.where(bits, t, 1 - t)))
                if wt > 1e-12:
                    idx.append(int(flat_index(c, *(i0 + bits), shape)))
                    w.append(wt * pdir[c])
        self.idx, self.w = np.asarray(idx, dtype=np.int64), np.asarray(w, dtype=float)
        self.coef = dt * np.asarray(inv_eps, dtype=float).reshape(-1)[self.idx] * self.w
        self.tau = 2.0 / omega_half
        self.t0 = 5.0 * self.tau
        t = np.arange(max_steps + 3) * dt
        g = np.exp(-(((t - self.t0) / self.tau) ** 2)) * np.sin(omega_c * (t - self.t0))
        self.wave = np.diff(g) / (omega_c * dt)
        self.dt = dt

    def bind(self, eng):
        self._coef = self.coef.astype(eng.rdt)
        self._wave = self.wave.astype(eng.rdt)

    def correct_H(self, eng):
        pass

    def advance_H(self, eng):
        pass

    def correct_E(self, eng):
        eng.E_flat[self.idx] -= self._coef * self._wave[eng.n]

    def advance_E(self, eng):
        pass


# --- TFSF box geometry (tfsf.py:25-105) -------------------------------------------------------

def _offsets(kind: str, c: int) -> np.ndarray:
    o = np.zeros(3)
    if kind == "E":
        o[c] = 0.5
    else:
        o[:] = 0.5
        o[c] = 0.0
    return o


def matched_eps_1d(khat, omega, dt, dx, eps_b) -> float:
    """1D permittivity whose Yee phase velocity matches the 3D grid's along khat (tfsf.py:35-47)."""
    khat = np.asarray(khat, dtype=float)
    lhs = eps_b * (np.sin(omega * dt / 2) / dt) ** 2
    lo, hi = 0.0, np.pi / dx
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        val = np.sum((np.sin(mid * khat * dx / 2) / dx) ** 2) - lhs
        lo, hi = (mid, hi) if val < 0 else (lo, mid)
    k = 0.5 * (lo + hi)
    return float((np.sin(k * dx / 2) / dx) ** 2 / (np.sin(omega * dt / 2) / dt) ** 2)


def tfsf_slabs(shape, lo, hi, dt, spacing, periodic=(False, False, False)):
    """(kind, c, region, g, coef, q) per one-cell correction slab of the box [lo, hi] (tfsf.py:50-105)."""
    spacing = np.broadcast_to(np.asarray(spacing, dtype=float), (3,))
    lo, hi = np.asarray(lo, dtype=float), np.asarray(hi, dtype=float)
    out = []
    for kind in ("E", "H"):
        for c in range(3):
            R = e_region(c, shape, periodic) if kind == "E" else h_region(c, shape, periodic)
            off = _offsets(kind, c)
            d1, d2 = (c + 1) % 3, (c + 2) % 3
            for d, g, s in ((d1, d2, 1.0), (d2, d1, -1.0)):
                if kind == "H":
                    s = -s
                trans = {}
                for a in range(3):
                    if a == d:
                        continue
                    ix = np.arange(R[a].start, R[a].stop)
                    ok = (ix + off[a] >= lo[a]) & (ix + off[a] <= hi[a])
                    trans[a] = ix[ok]
                if any(len(v) == 0 for v in trans.values()):
                    continue
                idx_d = np.arange(R[d].start, R[d].stop)
                u = idx_d + off[d]
                inside = lambda x, d=d: ((x >= lo[d]) & (x <= hi[d])).astype(int)  # noqa: E731
                for sign_q, du in ((1.0, 0.5), (-1.0, -0.5)):
                    delta = inside(u) - inside(u + du)
                    for i in np.nonzero(delta)[0]:
                        region, axes = [None] * 3, []
                        for a in range(3):
                            if a == d:
                                region[a] = slice(int(idx_d[i]), int(idx_d[i]) + 1)
                                axes.append(np.array([u[i] + du]))
                            else:
                                region[a] = slice(int(trans[a][0]), int(trans[a][-1]) + 1)
                                axes.append(trans[a] + off[a])
                        q = np.stack(np.meshgrid(*axes, indexing="ij"), axis=-1)
                        coef = s * sign_q * delta[i] * (dt / spacing[d])
                        out.append((kind, c, tuple(region), g, coef, q))
    return out


class TFSFPlaneWave:
    """TFSF box fed by an auxiliary 1D Yee grid (tfsf.py:108-211)."""

    def __init__(self, *, shape, origin, dx, dt, r0, lo, hi, khat, ehat, eps_b, omega_c, omega_half, max_steps,
                 periodic=(False, False, False), m0=None):
        spacing = np.broadcast_to(np.asarray(dx, dtype=float), (3,)).copy()
        self.shape, self.dt = tuple(shape), float(dt)
        self.khat, self.ehat = np.asarray(khat, float), np.asarray(ehat, float)
        self.hhat = np.cross(self.khat, self.ehat)
        self.eps_b = float(eps_b)
        self.axis_aligned = int(np.sum(np.abs(self.khat) > 1e-12)) == 1
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
        m_max = self.m0 + int(np.ceil(s_all.max() / ds)) + 4
        self.M = m_max + PML1D + 2
        self.slabs = []
        for kind, c, region, coef, sq in raw:
            f = sq.ravel() / ds + self.m0 - (0.5 if kind == "E" else 0.0)
            i0 = np.floor(f + 1e-9).astype(np.int64)
            w = f - i0
            w[np.abs(w) < 1e-9] = 0.0
            shape_ = tuple(r.stop - r.start for r in region)
            self.slabs.append((kind, c, region, float(coef), i0, w if np.any(w) else None, shape_))
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
        t = (np.arange(max_steps + 2) + 1) * dt
        self.wave = np.exp(-(((t - self.t0) / self.tau) ** 2)) * np.sin(omega_c * (t - self.t0))

    def probe(self):
        return np.array([self.m0]), np.array([self.m0 - 1]), np.array([self.m0])

    def bind(self, eng):
        dt_ = eng.rdt
        self.E_flat = np.zeros(self.M, dt_)
        self.H_flat = np.zeros(self.M, dt_)
        self._ca, self._cb = self.ca[1:-1].astype(dt_), self.cb[1:-1].astype(dt_)
        self._da, self._db = self.da[:-1].astype(dt_), self.db[:-1].astype(dt_)
        self._wave = self.wave.astype(dt_)
        self._slabs = [(k, c, r, coef, i0, None if w is None else w.astype(dt_), s)
                       for k, c, r, coef, i0, w, s in self.slabs]

    def _apply(self, eng, kind):
        target = eng.E if kind == "E" else eng.H
        src = self.H_flat if kind == "E" else self.E_flat
        for k, c, region, coef, i0, w, shape_ in self._slabs:
            if k != kind:
                continue
            vals = src[i0]
            if w is not None:
                vals = vals + w * (src[i0 + 1] - vals)
            target[c][region] += coef * vals.reshape(shape_)

    def correct_H(self, eng):
        self._apply(eng, "H")

    def correct_E(self, eng):
        self._apply(eng, "E")

    def advance_H(self, eng):
        E, H = self.E_flat, self.H_flat
        H[:-1] *= self._da
        H[:-1] -= self._db * (E[1:] - E[:-1])

    def advance_E(self, eng):
        E, H = self.E_flat, self.H_flat
        E[1:-1] *= self._ca
        E[1:-1] -= self._cb * (H[1:-1] - H[:-2])
        E[self.m_src] += self._wave[eng.n]


# --- table-fed TFSF, per step (layered.py:164-242, 472-482) ------------------------------------

STENCIL = np.arange(-2, 4)
DENOM = [float(np.prod([xj - xk for xk in STENCIL if xk != xj])) for xj in STENCIL]


def read_tables(T, n, m, cmax, base, i0, u0):
    """Quintic interpolation of tables sampled every m steps (layered.py:169-187)."""
    q = n + np.asarray(i0, dtype=np.int64)
    r = q % m
    c = np.minimum((q - r) // m, cmax)
    u = (r.astype(np.asarray(u0).dtype) + u0) / m
    d = [u - float(x) for x in STENCIL]
    idx = np.asarray(base, dtype=np.int64) + c
    out = None
    for j in range(len(STENCIL)):
        w = None
        for k in range(len(STENCIL)):
            if k != j:
                w = d[k] if w is None else w * d[k]
        term = (w / DENOM[j]) * T[idx + j]
        out = term if out is None else out + term
    return out


class TableSource:
    """Per-step part of a table-fed TFSF box: every correction point reads its own table row at its
    own delay and adds coef * value to its node (layered.py:190-242). ``points[kind]`` =
    (flat node index, base, i0, u0, coef); ``inc`` = (base, i0, u0) of the incident-wave view."""

    def __init__(self, T, m, cmax, points, inc=None, T_inc=None):
        self.T, self.m, self.cmax = np.asarray(T, dtype=float), int(m), int(cmax)
        self.points = points
        self.inc, self.T_inc = inc, T_inc

    def bind(self, eng):
        self._T = self.T.astype(eng.rdt)
        self._pts = {k: (np.asarray(v[0]), np.asarray(v[1]), np.asarray(v[2]), np.asarray(v[3], dtype=eng.rdt),
                         np.asarray(v[4], dtype=eng.rdt)) for k, v in self.points.items()}
        if self.inc is not None:
            self.E_flat = np.zeros(3, eng.rdt)
            self.H_flat = self.E_flat
            self._Tinc = np.asarray(self.T_inc).astype(eng.rdt)

    def _apply(self, eng, kind):
        if kind not in self._pts:
            return
        nodes, base, i0, u0, coef = self._pts[kind]
        vals = coef * read_tables(self._T, eng.n, self.m, self.cmax, base, i0, u0)
        target = eng.E_flat if kind == "E" else eng.H_flat
        np.add.at(target, nodes, vals)

    def correct_H(self, eng):
        self._apply(eng, "H")

    def correct_E(self, eng):
        self._apply(eng, "E")

    def advance_H(self, eng):
        pass

    def advance_E(self, eng):
        if self.inc is not None:
            base, i0, u0 = self.inc
            self.E_flat[...] = read_tables(self._Tinc, eng.n, self.m, self.cmax, base, i0,
                                           np.asarray(u0, dtype=eng.rdt))


# --- Bloch plane source (tfsf.py:214-296) -----------------------------------------------------

class BlochPlaneSource:
    def __init__(self, *, shape, spacing, dt, k_src, k_launch, kpar, pol, eps_b, pml, omega_c, omega_half,
                 max_steps):
        self.shape, self.spacing, self.dt = tuple(shape), tuple(float(d) for d in spacing), float(dt)
        self.k_src, self.k_launch = int(k_src), int(k_launch)
        self.kpar = (float(kpar[0]), float(kpar[1]))
        self.eps_b, self.pml = float(eps_b), pml
        dx, dy = self.spacing[0], self.spacing[1]
        K = np.array([2 * np.sin(self.kpar[a] * self.spacing[a] / 2) / self.spacing[a] for a in (0, 1)])
        u = K / np.linalg.norm(K)
        a = np.array([-u[1], u[0]]) if pol == "s" else u
        self.amp = (a[0] * np.exp(0.5j * self.kpar[0] * dx), a[1] * np.exp(0.5j * self.kpar[1] * dy))
        nx, ny = shape[0] - 2, shape[1] - 2
        i, j = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
        self.phase = np.exp(1j * (self.kpar[0] * i * dx + self.kpar[1] * j * dy))[:, :, None]
        self.tau = 2.0 / omega_half
        self.t0 = 3.5 * self.tau
        t = (np.arange(max_steps + 2) + 1) * dt
        self.wave = np.exp(-(((t - self.t0) / self.tau) ** 2)) * np.sin(omega_c * (t - self.t0))

    def probe(self, k):
        n = (3, 3, self.shape[2])
        e = np.array([flat_index(0, 1, 1, k, n), flat_index(1, 1, 1, k, n)])
        h1 = np.array([flat_index(1, 1, 1, k - 1, n), flat_index(0, 1, 1, k - 1, n)])
        h2 = np.array([flat_index(1, 1, 1, k, n), flat_index(0, 1, 1, k, n)])
        return e, h1, h2

    def bind(self, eng):
        nz = self.shape[2]
        self.col = Engine(shape=(3, 3, nz), dx=self.spacing, dt=self.dt, inv_eps=np.full((3, 3, 3, nz), 1 / self.eps_b),
                          pml=self.pml, periodic=(True, True, False), bloch=(*self.kpar, 0.0), dtype=eng.rdt)
        self._phase = self.phase.astype(eng.cdt)
        self._amp = [np.asarray(v, dtype=eng.cdt) for v in self.amp]
        self._wave = self.wave.astype(eng.rdt)
        nx, ny = self.shape[0] - 2, self.shape[1] - 2
        self._xy = (slice(1, nx + 1), slice(1, ny + 1))
        self._cE = self.dt / (self.eps_b * self.spacing[2])
        self._cH = self.dt / self.spacing[2]

    @property
    def E_flat(self):
        return self.col.E_flat

    @property
    def H_flat(self):
        return self.col.H_flat

    def _inc(self, F, c, k):
        return self._phase * F[c][1:2, 1:2, k:k + 1]

    def correct_H(self, eng):
        k, (X, Y) = self.k_src, self._xy
        E = self.col.E
        eng.H[0][X, Y, k - 1:k] -= self._cH * self._inc(E, 1, k)
        eng.H[1][X, Y, k - 1:k] += self._cH * self._inc(E, 0, k)

    def correct_E(self, eng):
        k, (X, Y) = self.k_src, self._xy
        H = self.col.H
        eng.E[0][X, Y, k:k + 1] += self._cE * self._inc(H, 1, k - 1)
        eng.E[1][X, Y, k:k + 1] -= self._cE * self._inc(H, 0, k - 1)

    def advance_H(self, eng):
        self.col.update_H()

    def advance_E(self, eng):
        self.col.update_E()
        w = self._wave[eng.n]
        k = self.k_launch
        self.col.E[0][1:2, 1:2, k:k + 1] += self._amp[0] * w
        self.col.E[1][1:2, 1:2, k:k + 1] += self._amp[1] * w


# --- mode source, per step (ports.py:362-372) --------------------------------------------------

class ModeInjection:
    """H_flat[h] += sum_r mH[:, r] tE[n, r]; E_flat[e] += sum_r mE[:, r] tH[n, r]."""

    def __init__(self, h_idx, mH, tE, e_idx, mE, tH):
        self.h_idx, self.mH, self.tE = np.asarray(h_idx), np.asarray(mH, float), np.asarray(tE, float)
        self.e_idx, self.mE, self.tH = np.asarray(e_idx), np.asarray(mE, float), np.asarray(tH, float)

    def bind(self, eng):
        self._mH, self._tE = self.mH.astype(eng.rdt), self.tE.astype(eng.rdt)
        self._mE, self._tH = self.mE.astype(eng.rdt), self.tH.astype(eng.rdt)

    def correct_H(self, eng):
        eng.H_flat[self.h_idx] += np.sum(self._mH * self._tE[eng.n], axis=1)

    def correct_E(self, eng):
        eng.E_flat[self.e_idx] += np.sum(self._mE * self._tH[eng.n], axis=1)

    def advance_H(self, eng):
        pass

    def advance_E(self, eng):
        pass
