"""Source tables of a table-fed TFSF plane wave: the grid's exact discrete plane wave at any angle, F_c(z, tau),
built on the host once per run (the construction half of ObliqueLayeredPlaneWave, layered.py:254-405 of the
reference; the per-step half is ``sources.TableSource`` and the HIP table kernels).

A plane wave at a fixed angle has the transverse slowness s = n sin(theta) (cos phi, sin phi) at every frequency,
so its field is F(z, t - s . r_par). The 1D problem runs along the axis with the largest |k| component; the grid
is relabelled by the cyclic permutation that takes that axis to z (which keeps the Yee staggering). Per frequency
the Yee equations for fields ~ exp(i w s . r_par) see transverse differences as i K_x, i K_y with
K = 2 sin(w s d / 2) / d, and the leapfrog step as Omega = 2 sin(w dt / 2) / dt (lightsmith's exp(-i w t)).

In a homogeneous host (the scattering problems here: no substrate) the background is the incident wave alone, and
the discrete wave has a closed form, used here instead of a transfer-matrix sweep: every component's complex
amplitude, referred to its own staggered position, obeys the continuum relations with k -> K, w -> Omega:

    Omega^2 eps = Kx^2 + Ky^2 + Kz^2,   K . E = 0,   H = K x E / Omega,   Kz = 2 sin(kz dz / 2) / dz,

so F_c(z) = A_c exp(i kz (z_c - z_end)) with the tangential E at the incidence-side end node equal to
ehat_t. A pulse g(t) (Gaussian envelope times sin(omega_c t), high-passed below omega_min / 3) is transformed,
multiplied by each component's response and transformed back; tables are kept every m steps (w h <= 0.2 at
omega_c + 3 omega_half), padded with zeros, and read per correction point with quintic interpolation at its own
delay. The time axis, pulse timing and padding follow layered.py:331-405 exactly, so the same box gets the same
source as the reference's.
"""

from __future__ import annotations

import numpy as np

from .media import Medium
from .sources import TableSource, tfsf_slabs

OMEGA_H = 0.2  # table sample spacing: omega h at omega_c + 3 omega_half (layered.py:166)
STENCIL_LO, STENCIL_N = -2, 6  # quintic stencil -2 .. 3 (layered.py:164)
NAMES = ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz")


class ObliqueLayeredPlaneWave(TableSource):
    """TFSF box [lo, hi] (node-index units) fed with the grid's exact discrete plane wave along khat, polarised
    along ehat, in the homogeneous host medium (layered.py:254-405 constructor contract; substrates, graded
    grids, mirror planes and beam foci are not supported here and raise).

    After construction: ``t_passed`` (the incident wave has passed every point, tables below 1e-11 of their
    peak), ``t_off`` (below 1e-6), ``n_inc`` (the host index), ``m`` (table stride in steps), ``view`` (the
    incident wave alone at the reference point, E_x, E_y, E_z, for a DFTMonitor's ``source``), ``background``
    and ``incident_reference`` (the frequency-domain background for flux bookkeeping)."""

    def __init__(self, *, shape, origin, spacing, dt, lo, hi, khat, ehat, inv_eps, host: Medium,
                 z_int: float = -np.inf, sub: Medium | None = None, omega_c, omega_half, max_steps: int,
                 omega_min: float | None = None, focus=None, t_focus: float = 0.0, mirror=(0, 0, 0), grid=None):
        if sub is not None:
            raise ValueError("layered (substrate) backgrounds are not supported by this source; the host is uniform")
        if grid is not None:
            raise ValueError("graded grids are not supported (grid='uniform')")
        if any(mirror):
            raise ValueError("mirror planes are not supported (symmetry='off')")
        if focus is not None:
            raise ValueError("beam foci are not supported (plane waves only)")
        if host.dispersive:
            raise ValueError("oblique incidence needs a non-dispersive incidence medium")
        khat = np.asarray(khat, dtype=float) / np.linalg.norm(khat)
        ehat = np.asarray(ehat, dtype=float) / np.linalg.norm(ehat)
        a = int(np.argmax(np.abs(khat)))
        self.perm = perm = np.array([(j + a + 1) % 3 for j in range(3)])  # relabelled axis j is grid axis perm[j]
        self.inv = inv = np.argsort(perm)
        self.shape, self.dt = tuple(int(v) for v in shape), float(dt)
        sp_grid = np.broadcast_to(np.asarray(spacing, dtype=float), (3,)).copy()
        self.spacing = sp = sp_grid[perm]
        self.origin = np.asarray(origin, dtype=float)[perm]
        self.host = host
        self.eps = float(host.eps_inf)
        self.n_inc = float(np.sqrt(self.eps))
        lo_g, hi_g = np.asarray(lo), np.asarray(hi)
        lo_r, hi_r = lo_g[perm], hi_g[perm]
        self.khat, self.ehat = khat[perm], ehat[perm]
        self.down = bool(self.khat[2] < 0)
        self.s = s = self.n_inc * self.khat[:2]
        dz = sp[2]
        self.k0 = int(lo_r[2]) - 2
        self.nk = nk = int(hi_r[2]) + 2 - self.k0 + 1
        self.z_nodes = self.origin[2] + (self.k0 + np.arange(nk)) * dz
        # the box corner the wave reaches first: every delay s . (r_par - r_c) >= 0
        corners = np.array([[x, y] for x in (lo_r[0], hi_r[0]) for y in (lo_r[1], hi_r[1])], dtype=float)
        corners = self._par(corners)
        self.r_c = corners[np.argmin(corners @ s)]
        self.z_end = self.z_nodes[-1] if self.down else self.z_nodes[0]
        self.r_ref = np.array([*self.r_c, self.z_end])[inv]
        self.m = m = max(1, int(OMEGA_H / ((omega_c + 3 * omega_half) * dt)))

        inv_eps = np.asarray(inv_eps, dtype=float)
        slabs = []
        for kind, c, region, g, coef, q in tfsf_slabs(self.shape, lo_g, hi_g, dt, sp_grid):
            comp = ("H" if kind == "E" else "E") + "xyz"[inv[g]]
            qr = q[..., perm]
            row = np.floor(qr[..., 2] + 1e-9).astype(np.int64) - self.k0
            delay = (self._par(qr[..., :2]) - self.r_c) @ s
            ext = tuple(r.stop - r.start for r in region)
            w = coef * (inv_eps[c][region] if kind == "E" else np.ones(ext))
            slabs.append((kind, c, region, comp, row.reshape(-1), delay.reshape(-1), w.reshape(-1)))
        delay_max = max(float(sl[5].max()) for sl in slabs)
        self.tau_lead = delay_max + 6 * m * dt
        self.tau_p = 2.0 / omega_half
        # the pulse starts at e^-25 of its peak; the high-pass below the band spreads it by ~1 / omega_lo
        self.omega_lo = omega_min / 3 if omega_min else None
        self.t0 = max(5.0 * self.tau_p, 10.0 / self.omega_lo if self.omega_lo else 0.0)
        span = self.tau_lead + 2 * self.t0 + 4 * nk * dz * self.n_inc + 4 * self.tau_p
        N = 1 << int(np.ceil(np.log2(span / dt)))
        while True:
            table, inc_table = self._transform(N, omega_c, omega_half)
            if np.abs(table[..., -N // 16:]).max() < 1e-6 * np.abs(table).max():
                break
            if N >= 1 << 20:
                raise ValueError("table-fed plane wave: the background does not decay within the table")
            N *= 2
        self.N = N
        # t_passed: every point has seen the tables fall below 1e-11 of their peak; t_off: below 1e-6. The tail is
        # searched between the peak and the quietest sample after it (past that, the periodic inverse transform
        # holds the wrapped leading edge)
        env = np.abs(table).reshape(-1, N).max(axis=0)
        i_pk = int(np.argmax(env))
        i_q = i_pk + int(np.argmin(env[i_pk:]))
        tail = env[i_pk:i_q + 1]
        i_off = i_pk + int(np.nonzero(tail > 1e-6 * env[i_pk])[0][-1])
        i_end = i_pk + int(np.nonzero(tail > 1e-11 * env[i_pk])[0][-1])
        self.t_passed = i_end * dt - self.tau_lead + delay_max
        self.t_off = i_off * dt - self.tau_lead + delay_max
        table, inc_table = table[..., ::m], inc_table[..., ::m]
        n_c = table.shape[-1]
        pad = STENCIL_N + 2  # zero samples after each row: reads clamped at cmax = n_c + 2 give 0
        L = n_c + pad
        cmax = n_c + 2
        T = np.concatenate([table, np.zeros(table.shape[:2] + (pad,))], axis=-1)
        T_inc = np.concatenate([inc_table, np.zeros((3, pad))], axis=-1)
        # per correction point: first stencil sample of its row (before the step's own offset), and its fine
        # position at n = 0 split into an integer and a fraction (E corrections read H at n + 1/2)
        points = {}
        for kind in ("H", "E"):
            parts = [sl for sl in slabs if sl[0] == kind]
            if not parts:
                continue
            nodes, base, i0, u0, w = [], [], [], [], []
            for _, c, region, comp, row, delay, wt in parts:
                f0 = (self.tau_lead - delay) / dt + (0.5 if kind == "E" else 0.0)
                ii = np.floor(f0).astype(np.int64)
                nodes.append(_region_nodes(self.shape, c, region))
                base.append((NAMES.index(comp) * nk + row) * L + STENCIL_LO)
                i0.append(ii)
                u0.append(f0 - ii)
                w.append(wt)
            points[kind] = tuple(np.concatenate(v) for v in (nodes, base, i0, u0, w))
        f0 = self.tau_lead / dt + 1  # the view holds E at the end of the step
        i0 = int(np.floor(f0))
        inc = (np.arange(3) * L + STENCIL_LO, np.full(3, i0), np.full(3, f0 - i0))
        super().__init__(T, m, cmax, points, inc=inc, T_inc=T_inc)

    # --- geometry -----------------------------------------------------------------------------------------
    def _par(self, u):
        """World transverse positions (..., 2) of relabelled fractional indices u (..., 2)."""
        return self.origin[:2] + np.asarray(u) * self.spacing[:2]

    # --- the discrete plane wave ---------------------------------------------------------------------------
    def _wave(self, w):
        """Component amplitudes (n_w, 6) (E_x, E_y, E_z, H_x, H_y, H_z in the relabelled frame, each referred to
        its own staggered position) and kz (n_w,) of the discrete plane wave whose tangential E is ehat_t."""
        w = np.asarray(w, dtype=float)
        sp, dz = self.spacing, self.spacing[2]
        Kx = 2 * np.sin(w * self.s[0] * sp[0] / 2) / sp[0]
        Ky = 2 * np.sin(w * self.s[1] * sp[1] / 2) / sp[1]
        Om = 2 * np.sin(w * self.dt / 2) / self.dt
        Kz = np.sqrt((Om ** 2 * self.eps - Kx ** 2 - Ky ** 2).astype(complex))
        if self.down:
            Kz = -Kz
        kz = 2 / dz * np.arcsin(Kz * dz / 2)
        ex = np.full(w.shape, self.ehat[0], dtype=complex)
        ey = np.full(w.shape, self.ehat[1], dtype=complex)
        ez = -(Kx * ex + Ky * ey) / Kz
        hx = (Ky * ez - Kz * ey) / Om
        hy = (Kz * ex - Kx * ez) / Om
        hz = (Kx * ey - Ky * ex) / Om
        return np.stack([ex, ey, ez, hx, hy, hz], axis=1), kz

    def _response(self, w) -> dict:
        """The discrete wave on the table rows: (n_w, nk) per component (E_x, E_y, H_z at z_k; E_z, H_x, H_y at
        z_k + dz / 2), and the incident wave at the end node, ``inc`` (n_w, 3): E_x, E_y (node), E_z (node + 1/2)."""
        A, kz = self._wave(w)
        dz = self.spacing[2]
        z_node = self.z_nodes - self.z_end
        out = {}
        for i, nm in enumerate(NAMES):
            half = nm in ("Ez", "Hx", "Hy")
            out[nm] = A[:, i:i + 1] * np.exp(1j * kz[:, None] * (z_node + (dz / 2 if half else 0.0))[None, :])
        out["inc"] = np.stack([A[:, 0], A[:, 1], A[:, 2] * np.exp(0.5j * kz * dz)], axis=1)
        return out

    def _transform(self, N: int, omega_c: float, omega_half: float):
        """Tables (6, nk, N) and the incident view's (3, N) over tau = j dt - tau_lead, j < N."""
        dt = self.dt
        tau = np.arange(N) * dt - self.tau_lead
        t_c = tau - self.t0
        X = np.fft.rfft(np.exp(-((t_c / self.tau_p) ** 2)) * np.sin(omega_c * t_c))
        om = 2 * np.pi * np.arange(len(X)) / (N * dt)
        if self.omega_lo:
            X = X * (1 - np.exp(-((om / self.omega_lo) ** 2)))
        keep = (om > 0) & (om < omega_c + 6 * omega_half)  # past the pulse band the grid response means nothing
        f = self._response(om[keep])
        spec = np.zeros((6, self.nk, len(X)), dtype=complex)
        for i, nm in enumerate(NAMES):
            # numpy's inverse transform is exp(+i w t): the exp(-i w t) response enters conjugated
            spec[i][:, keep] = np.conj(f[nm].T) * X[keep][None, :]
        inc = np.zeros((3, len(X)), dtype=complex)
        inc[:, keep] = np.conj(f["inc"].T) * X[keep][None, :]
        return np.fft.irfft(spec, n=N, axis=-1), np.fft.irfft(inc, n=N, axis=-1)[self.inv]

    # --- frequency-domain background (flux bookkeeping) ----------------------------------------------------
    def background(self, omega, flat, kind: str = "E") -> np.ndarray:
        """Background E (kind "E") or H at reference flat grid indices, (n_omega, n), relative to the incident
        wave whose tangential E at the reference point is ehat_t (layered.py:421-443)."""
        n_cells = int(np.prod(self.shape))
        comp, rest = np.divmod(np.asarray(flat, dtype=np.int64), n_cells)
        idx3 = np.stack(np.unravel_index(rest, self.shape), axis=1).astype(float)
        off = np.zeros((len(comp), 3))
        if kind == "E":
            off[np.arange(len(comp)), comp] = 0.5
        else:
            off[:] = 0.5
            off[np.arange(len(comp)), comp] = 0.0
        pos = (idx3 + off)[:, self.perm]  # relabelled fractional indices
        rc = self.inv[comp]  # relabelled component
        A, kz = self._wave(omega)
        z = self.origin[2] + pos[:, 2] * self.spacing[2] - self.z_end
        r_par = self._par(pos[:, :2])
        w = np.asarray(omega, dtype=float)
        phase = np.exp(1j * (w[:, None] * ((r_par - self.r_c) @ self.s)[None, :] + kz[:, None] * z[None, :]))
        col = rc + (3 if kind == "H" else 0)
        return A[:, col] * phase

    def incident_reference(self, omega) -> np.ndarray:
        """The incident wave alone at the reference point, (n_omega, 3) in grid components."""
        return self._response(np.asarray(omega, dtype=float))["inc"][:, self.inv]

    @property
    def view(self):
        """The monitor source of the incident wave at the reference point (E_x, E_y, E_z at indices 0..2)."""
        return self


def _region_nodes(shape, c: int, region) -> np.ndarray:
    from .engine import flat_index
    ii, jj, kk = np.meshgrid(*[np.arange(r.start, r.stop) for r in region], indexing="ij")
    return flat_index(c, ii, jj, kk, shape).reshape(-1)


__all__ = ["OMEGA_H", "ObliqueLayeredPlaneWave"]
