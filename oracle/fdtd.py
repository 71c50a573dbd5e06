"""FP64 restatement of lightsmith's FDTD ``Engine`` (oracle; test infrastructure only).

Follows ``engine.py`` of the reference for a uniform grid (cubic, or per-axis spacings) without
mirror planes: fields E, H of shape (3, N0, N1, N2) in C order; E_c at node + e_c / 2 and H_c at
node + (1 - e_c) / 2; the outermost node planes are PEC; periodic axes carry one ghost plane at
each end. ``dtype`` selects float64 (the oracle proper) or float32 (the reference's single-precision
op order, for measuring how much float32 rounding alone moves the fields).
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from .media import Medium, pole_coefficients


def flat_index(c, i, j, k, shape):
    """Reference flat index into the stacked (3, N0, N1, N2) array (engine.py:103-105)."""
    n0, n1, n2 = shape
    return ((np.asarray(c) * n0 + np.asarray(i)) * n1 + np.asarray(j)) * n2 + np.asarray(k)


def e_region(c: int, shape, periodic=(False, False, False)):
    """Stepped index box of E_c: tangential E on the outer planes stays 0 (engine.py:47-60, no mirror)."""
    return tuple(slice(1, n - 1) if (periodic[a] or a != c) else slice(0, n - 1) for a, n in enumerate(shape))


def h_region(c: int, shape, periodic=(False, False, False)):
    """Stepped index box of H_c (engine.py:63-66, no mirror)."""
    return tuple(slice(1, n - 1) if periodic[a] else (slice(0, n) if a == c else slice(0, n - 1))
                 for a, n in enumerate(shape))


def _moved(R, d: int, s: int):
    R = list(R)
    R[d] = slice(R[d].start + s, R[d].stop + s)
    return tuple(R)


@dataclass
class PMLSpec:  # engine.py:184-191
    cells: tuple
    m: float = 3.0
    kappa_max: float = 1.0
    alpha_max: float = 0.0
    eps_ref: float = 1.0
    eps_ref_high: float | None = None


def cpml_profile(spec: PMLSpec, n_lo: int, n_hi: int, N: int, h: float, dt: float, u):
    """(b, c, kappa) of the CPML at positions u (cell units) along an axis of N nodes (engine.py:358-371)."""
    u = np.asarray(u, dtype=float)
    s_lo = 0.8 * (spec.m + 1) / (np.sqrt(spec.eps_ref) * h)
    s_hi = 0.8 * (spec.m + 1) / (np.sqrt(spec.eps_ref_high or spec.eps_ref) * h)
    r_lo = np.clip((n_lo - u) / n_lo, 0, 1) if n_lo else np.zeros_like(u)
    r_hi = np.clip((u - (N - 1 - n_hi)) / n_hi, 0, 1) if n_hi else np.zeros_like(u)
    rho = np.maximum(r_lo, r_hi)
    sig = np.where(r_lo >= r_hi, s_lo, s_hi) * rho ** spec.m
    kap = 1 + (spec.kappa_max - 1) * rho ** spec.m
    alp = spec.alpha_max * (1 - rho) * (rho > 0)
    b = np.exp(-(sig / kap + alp) * dt)
    den = sig * kap + kap ** 2 * alp
    c = np.where(sig > 0, sig * (b - 1) / np.where(den > 0, den, 1), 0.0)
    return b, c, kap


def cpml_ranges(start: int, stop: int, n_lo: int, n_hi: int, N: int):
    """Index ranges along the PML axis of one (field, component) region (engine.py:388-395)."""
    out = []
    if n_lo:
        out.append((start, min(stop, n_lo + 1)))
    if n_hi:
        out.append((max(start, N - 2 - n_hi), stop))
    return [(a, b) for a, b in out if b > a]


@dataclass
class DispersiveGroup:  # engine.py:129-137
    idx: np.ndarray
    media: list
    frac: np.ndarray
    w: np.ndarray | None


@dataclass
class OffDiagonal:  # engine.py:117-126
    idx3: np.ndarray
    comp: np.ndarray
    coef: np.ndarray


@dataclass
class DFTMonitor:  # engine.py:194-210
    omega: np.ndarray
    e_idx: np.ndarray
    h_idx1: np.ndarray | None = None
    h_idx2: np.ndarray | None = None
    energy: int = 0
    source: object | None = None
    acc_e: np.ndarray | None = field(default=None, repr=False)
    acc_h: np.ndarray | None = field(default=None, repr=False)

    def results(self):
        return self.acc_e.astype(complex), None if self.acc_h is None else self.acc_h.astype(complex)


@dataclass
class TimeMonitor:  # engine.py:213-224
    e_idx: np.ndarray
    n_max: int
    buf: np.ndarray | None = field(default=None, repr=False)
    pos: int = 0

    def results(self, n: int):
        return self.buf[:n].T


class Engine:
    """The reference update, one half step at a time (engine.py:230-719)."""

    def __init__(self, *, shape, dx, dt, inv_eps, groups=(), pml=None, sources=(), monitors=(), dft_stride=1,
                 periodic=(False, False, False), bloch=(0.0, 0.0, 0.0), offdiag=None, time_monitors=(),
                 dtype=np.float64):
        self.shape = tuple(int(n) for n in shape)
        self.dt = float(dt)
        self.spacing = tuple(float(v) for v in np.broadcast_to(np.asarray(dx, dtype=float), (3,)))
        self.uniform = len(set(self.spacing)) == 1
        self.dx = self.spacing[0]
        self.periodic = tuple(bool(p) for p in periodic)
        self.bloch = tuple(float(k) for k in bloch)
        if any(k and not p for k, p in zip(self.bloch, self.periodic)):
            raise ValueError("a Bloch wavevector needs a periodic axis")
        self.complex = any(self.bloch)
        self.rdt = np.dtype(dtype)
        self.cdt = np.complex64 if self.rdt == np.float32 else np.complex128
        self.fdt = self.cdt if self.complex else self.rdt
        self.stride = int(dft_stride)
        self.E = np.zeros((3,) + self.shape, self.fdt)
        self.H = np.zeros((3,) + self.shape, self.fdt)
        self.n = 0
        cdx = self.dt / self.dx if self.uniform else self.dt  # engine.py:275
        self.inv_d = tuple(1.0 / d for d in self.spacing)
        self.RE = [e_region(c, self.shape, self.periodic) for c in range(3)]
        self.RH = [h_region(c, self.shape, self.periodic) for c in range(3)]
        inv_eps = np.asarray(inv_eps, dtype=float)
        self.cbE = [(cdx * inv_eps[c][self.RE[c]]).astype(self.rdt) for c in range(3)]  # engine.py:299
        self.cH = cdx
        self._build_pml(pml or PMLSpec(cells=((0, 0),) * 3), inv_eps)
        self._build_offdiag(offdiag, list(groups))
        self.groups = [_group(self, g) for g in groups]
        self.sources = list(sources)
        for s in self.sources:
            s.bind(self)
        self.monitors = list(monitors)
        for m in self.monitors:
            m.acc_e = np.zeros((len(m.omega), len(m.e_idx)), self.cdt)
            if m.h_idx1 is not None:
                m.acc_h = np.zeros((len(m.omega), len(m.h_idx1)), self.cdt)
        self.time_monitors = list(time_monitors)
        for m in self.time_monitors:
            m.buf = np.zeros((m.n_max, len(m.e_idx)), self.fdt)
            m.pos = 0
        self.energy = np.zeros(2, self.rdt)

    @property
    def E_flat(self):
        return self.E.reshape(-1)

    @property
    def H_flat(self):
        return self.H.reshape(-1)

    # --- CPML (engine.py:344-418) ---------------------------------------------------
    def _build_pml(self, spec: PMLSpec, inv_eps):
        self.pml_E, self.pml_H = [], []
        dt = self.dt
        for d in range(3):
            n_lo, n_hi = spec.cells[d]
            if not (n_lo or n_hi):
                continue
            if self.periodic[d]:
                raise ValueError("a periodic axis cannot have a PML")
            N, h = self.shape[d], self.spacing[d]
            for kind, regions in (("E", self.RE), ("H", self.RH)):
                for c in range(3):
                    if c == d:
                        continue
                    d1, d2 = (c + 1) % 3, (c + 2) % 3
                    src, sgn = (d2, 1.0) if d == d1 else (d1, -1.0)
                    if kind == "H":
                        sgn = -sgn
                    R = regions[c]
                    for a, b_ in cpml_ranges(R[d].start, R[d].stop, n_lo, n_hi, N):
                        Rs = list(R)
                        Rs[d] = slice(a, b_)
                        Rs = tuple(Rs)
                        u = np.arange(a, b_) + (0.5 if kind == "H" else 0.0)
                        bb, cc, kk = cpml_profile(spec, n_lo, n_hi, N, h, dt, u)
                        bshape = [1, 1, 1]
                        bshape[d] = b_ - a
                        cd = dt / h if self.uniform else dt
                        if kind == "E":
                            scale = cd * inv_eps[c][Rs]
                        else:
                            scale = np.full([s.stop - s.start for s in Rs], cd)
                        if not self.uniform:
                            scale = scale / h
                        cp = (sgn * scale * cc.reshape(bshape)).astype(self.rdt)
                        kp = (sgn * scale * (1 / kk - 1).reshape(bshape)).astype(self.rdt) if np.any(kk > 1) else None
                        term = (c, src, d, Rs, bb.reshape(bshape).astype(self.rdt), cp, kp,
                                np.zeros(cp.shape, self.fdt))
                        (self.pml_E if kind == "E" else self.pml_H).append(term)

    # --- off-diagonal tensor smoothing (engine.py:421-518, uniform grid, no mirror) ---------
    def _build_offdiag(self, od, groups):
        self.od = None
        if od is None or not len(od.comp):
            return
        shape = self.shape
        nflat = 3 * int(np.prod(shape))
        idx3 = np.asarray(od.idx3, dtype=np.int64)
        comp = np.asarray(od.comp, dtype=np.int64)
        coef = np.asarray(od.coef, dtype=float)
        tc, ti, sc, si, vals, back = [], [], [], [], [], []
        for d in range(3):
            for corner in range(4):
                q = idx3.copy()
                q += np.eye(3, dtype=np.int64)[comp] * (corner & 1)
                q[:, d] -= corner >> 1
                phase = np.ones(len(q), dtype=complex)
                ok = (comp != d) & (coef[:, d] != 0)
                for a in range(3):
                    n = shape[a]
                    if self.periodic[a]:
                        kl = self.bloch[a] * (n - 2) * self.spacing[a]
                        lo, hi = q[:, a] == 0, q[:, a] == n - 1
                        q[lo, a], q[hi, a] = n - 2, 1
                        phase[lo] *= np.exp(-1j * kl)
                        phase[hi] *= np.exp(1j * kl)
                    else:
                        rd = e_region(d, shape)[a]
                        ok &= (q[:, a] < rd.stop) & (q[:, a] >= rd.start)
                tc.append(comp[ok])
                ti.append(idx3[ok])
                sc.append(np.full(int(ok.sum()), d))
                si.append(q[ok])
                vals.append(coef[ok, d] / 8 * phase[ok])
                back.append(np.conj(coef[ok, d] * phase[ok]) / 8)
        tc, ti, sc, si, v = (np.concatenate(x) for x in (tc, ti, sc, si, vals))
        tc, sc = np.concatenate([tc, sc]), np.concatenate([sc, tc])
        ti, si = np.concatenate([ti, si]), np.concatenate([si, ti])
        v = np.concatenate([v, np.concatenate(back)])
        t = flat_index(tc, ti[:, 0], ti[:, 1], ti[:, 2], shape)
        s_ = flat_index(sc, si[:, 0], si[:, 1], si[:, 2], shape)
        if groups:
            disp = np.concatenate([np.asarray(g.idx) for g in groups])
            keep = ~(np.isin(t, disp) | np.isin(s_, disp))
            t, s_, v = t[keep], s_[keep], v[keep]
        if not len(t):
            return
        key, inv = np.unique(t * nflat + s_, return_inverse=True)
        v = np.bincount(inv, v.real, len(key)) + 1j * np.bincount(inv, v.imag, len(key))
        t, s_ = key // nflat, key % nflat
        targets, start, count = np.unique(t, return_index=True, return_counts=True)
        slot = np.arange(len(t)) - np.repeat(start, count)
        row = np.repeat(np.arange(len(targets)), count)
        width = int(count.max())
        h = np.ones((len(targets), 4 * width), dtype=np.int64)
        w = np.zeros((len(targets), 4 * width), dtype=complex)
        d, rest = np.divmod(s_, int(np.prod(shape)))
        q = np.stack(np.unravel_index(rest, shape), axis=1)
        d1, d2 = (d + 1) % 3, (d + 2) % 3
        inv_d = np.asarray(self.inv_d)
        eye = np.eye(3, dtype=np.int64)
        for k, (cmp, bk, sgn, ax) in enumerate(((d2, None, 1.0, d1), (d2, d1, -1.0, d1),
                                                (d1, None, -1.0, d2), (d1, d2, 1.0, d2))):
            p = q if bk is None else q - eye[bk]
            h[row, 4 * slot + k] = flat_index(cmp, p[:, 0], p[:, 1], p[:, 2], shape)
            w[row, 4 * slot + k] = self.dt * sgn * inv_d[ax] * v
        if not self.complex:
            w = w.real
        self.od = (targets, h, w.astype(self.fdt))

    # --- monitors (engine.py:564-592) ------------------------------------------------
    def _phase(self, omega, t):
        wt = np.asarray(omega, dtype=np.float64) * t
        return (np.cos(wt) + 1j * np.sin(wt)).astype(self.cdt)

    def sample(self):
        t_e = float(self.n) * self.dt
        t_h = t_e - 0.5 * self.dt
        energy = [None, None]
        for m in self.monitors:
            src = self if m.source is None else m.source
            ef, hf = src.E_flat, src.H_flat
            se = ef[np.asarray(m.e_idx)]
            m.acc_e += self._phase(m.omega, t_e)[:, None] * se[None, :]
            if m.acc_h is not None:
                sh = 0.5 * (hf[np.asarray(m.h_idx1)] + hf[np.asarray(m.h_idx2)])
                m.acc_h += self._phase(m.omega, t_h)[:, None] * sh[None, :]
                if m.energy:
                    if self.complex:
                        v = np.sum(np.real(se * np.conj(se))) + np.sum(np.real(sh * np.conj(sh)))
                    else:
                        v = np.sum(se * se) + np.sum(sh * sh)
                    g = int(m.energy) - 1
                    energy[g] = v if energy[g] is None else energy[g] + v
        for g, v in enumerate(energy):
            if v is not None:
                self.energy[g] = v
        for m in self.time_monitors:
            m.buf[m.pos] = self.E_flat[np.asarray(m.e_idx)]
            m.pos += 1

    # --- stepping (engine.py:595-683) -------------------------------------------------
    def ghosts(self, F, low: bool):
        for a, per in enumerate(self.periodic):
            if not per:
                continue
            n = self.shape[a]
            dst, src = (0, n - 2) if low else (n - 1, 1)
            ix_d, ix_s = [slice(None)] * 4, [slice(None)] * 4
            ix_d[a + 1], ix_s[a + 1] = slice(dst, dst + 1), slice(src, src + 1)
            k = self.bloch[a]
            if k:
                kl = k * (n - 2) * self.spacing[a] * (-1 if low else 1)
                F[tuple(ix_d)] = F[tuple(ix_s)] * complex(np.cos(kl), np.sin(kl))
            else:
                F[tuple(ix_d)] = F[tuple(ix_s)]

    def _curl(self, F, c, R, forward: bool, coef):
        d1, d2 = (c + 1) % 3, (c + 2) % 3
        if forward:
            a = F[d2][_moved(R, d1, 1)] - F[d2][R]
            b = F[d1][_moved(R, d2, 1)] - F[d1][R]
        else:
            a = F[d2][R] - F[d2][_moved(R, d1, -1)]
            b = F[d1][R] - F[d1][_moved(R, d2, -1)]
        if not self.uniform:
            a = a * self.inv_d[d1]
            b = b * self.inv_d[d2]
        return (a - b) * coef

    def update_H(self):
        E, H = self.E, self.H
        if any(self.periodic):
            self.ghosts(E, low=False)
        for c in range(3):
            H[c][self.RH[c]] -= self._curl(E, c, self.RH[c], True, self.cH)
        for c, src, d, Rs, b, cp, kp, psi in self.pml_H:
            der = E[src][_moved(Rs, d, 1)] - E[src][Rs]
            psi *= b
            if kp is not None:
                H[c][Rs] += kp * der
            psi += cp * der
            H[c][Rs] += psi

    def update_E(self):
        E, H = self.E, self.H
        if any(self.periodic):
            self.ghosts(H, low=True)
        for c in range(3):
            E[c][self.RE[c]] += self._curl(H, c, self.RE[c], False, self.cbE[c])
        if self.od is not None:
            e_idx, h_idx, w = self.od
            self.E_flat[e_idx] += np.sum(w * self.H_flat[h_idx], axis=1)
        for c, src, d, Rs, b, cp, kp, psi in self.pml_E:
            der = H[src][Rs] - H[src][_moved(Rs, d, -1)]
            psi *= b
            if kp is not None:
                E[c][Rs] += kp * der
            psi += cp * der
            E[c][Rs] += psi

    def step(self):
        self.update_H()
        for s in self.sources:
            s.correct_H(self)
        for s in self.sources:
            s.advance_H(self)
        self.update_E()
        for s in self.sources:
            s.correct_E(self)
        for g in self.groups:
            g.step(self)
        for s in self.sources:
            s.advance_E(self)
        self.n += 1

    def chunk(self):
        for _ in range(self.stride):
            self.step()
        self.sample()

    def run(self, steps: int):
        for _ in range(steps):
            self.step()


def _advance(pairs, coefs, e):
    """Advance every pole from e; return the sum of the new P (engine.py:759-767)."""
    tot = None
    for pair, (c1, c2, c3) in zip(pairs, coefs):
        new = c1 * pair[0] + c2 * pair[1] + c3 * e
        pair[1] = pair[0]
        pair[0] = new
        tot = new if tot is None else tot + new
    return tot


def _group(eng, g):
    """The group of a DispersiveGroup (pure or mixed) or of a TensorMesh (oracle/tensor.py)."""
    from .tensor import TensorGroup, TensorMesh
    if isinstance(g, TensorMesh):
        return TensorGroup(eng, g)
    return PureGroup(eng, g) if g.w is None else MixedGroup(eng, g)


class PureGroup:
    """One dispersive medium in D-form: D += E - E_prev; P from E_prev; E = (D - sum P) / eps_inf
    (engine.py:722-742, 769-793)."""

    def __init__(self, eng: Engine, g: DispersiveGroup):
        n = len(g.idx)
        self.idx = np.asarray(g.idx, dtype=np.int64)
        (m,) = g.media
        self.inv_eps_inf = 1.0 / m.eps_inf
        self.coef = [pole_coefficients(p, eng.dt) for p in m.poles]
        self.D = np.zeros(n, eng.fdt)
        self.E_prev = np.zeros(n, eng.fdt)
        self.P = [[np.zeros(n, eng.fdt), np.zeros(n, eng.fdt)] for _ in m.poles]

    def step(self, eng: Engine):
        e_now = eng.E_flat[self.idx]
        self.D += e_now - self.E_prev
        tot = _advance(self.P, self.coef, self.E_prev)
        e_new = (self.D - tot) * self.inv_eps_inf if tot is not None else self.D * self.inv_eps_inf
        eng.E_flat[self.idx] = e_new
        self.E_prev = e_new.copy()


class MixedGroup:
    """Interface cells with two parallel paths, w <1/eps> + (1 - w) / <eps> (engine.py:743-757, 775-791)."""

    def __init__(self, eng: Engine, g: DispersiveGroup):
        n = len(g.idx)
        dt = eng.dt
        self.idx = np.asarray(g.idx, dtype=np.int64)
        f = np.asarray(g.frac, dtype=float)
        media = list(g.media)
        self.coef = [[pole_coefficients(p, dt) for p in m.poles] for m in media]
        self.disp = [s for s, m in enumerate(media) if m.dispersive]
        self.h0 = sum(f[s] / m.eps_inf for s, m in enumerate(media) if not m.dispersive) + 0 * f[0]
        self.f = {s: f[s] for s in self.disp}
        self.inv_eps_inf = {s: 1.0 / media[s].eps_inf for s in self.disp}
        self.inv_eps_a = 1.0 / sum(f[s] * m.eps_inf for s, m in enumerate(media))
        self.w = np.asarray(g.w, dtype=float)
        self.has_h = bool(np.any(self.w > 0))
        self.has_a = bool(np.any(self.w < 1))
        z = lambda: np.zeros(n, eng.fdt)  # noqa: E731
        self.D, self.E_prev = z(), z()
        self.Eh = {s: z() for s in self.disp}
        self.Ea = z()
        self.Ph = {s: [[z(), z()] for _ in media[s].poles] for s in self.disp}
        self.Pa = {s: [[z(), z()] for _ in media[s].poles] for s in self.disp}

    def step(self, eng: Engine):
        e_now = eng.E_flat[self.idx]
        self.D += e_now - self.E_prev
        parts = []
        if self.has_h:
            harm = self.D * self.h0
            for s in self.disp:
                tot = _advance(self.Ph[s], self.coef[s], self.Eh[s])
                self.Eh[s] = (self.D - tot) * self.inv_eps_inf[s]
                harm = harm + self.f[s] * self.Eh[s]
            parts.append(self.w * harm)
        if self.has_a:
            acc = self.D
            for s in self.disp:
                tot = _advance(self.Pa[s], self.coef[s], self.Ea)
                acc = acc - self.f[s] * tot
            self.Ea = acc * self.inv_eps_a
            parts.append((1 - self.w) * self.Ea)
        e_new = parts[0] if len(parts) == 1 else parts[0] + parts[1]
        eng.E_flat[self.idx] = e_new
        self.E_prev = np.array(e_new, copy=True)


def medium_eps_inf(m: Medium) -> float:
    return float(m.eps_inf)
