"""Source tables of a table-fed TFSF plane wave: the grid's exact discrete plane wave at any angle, F_c(z, tau),
built on the host once per run (the construction half of ObliqueLayeredPlaneWave, layered.py:254-405 of the
reference; the per-step half is ``sources.TableSource`` and the HIP table kernels).

A plane wave at a fixed angle has the transverse slowness s = n sin(theta) (cos phi, sin phi) at every frequency,
so its field is F(z, t - s . r_par). The 1D problem runs along the axis with the largest |k| component; the grid
is relabelled by the cyclic permutation that takes that axis to z (which keeps the Yee staggering). Per frequency
the Yee equations for fields ~ exp(i w s . r_par) see transverse differences as i K_x, i K_y with
K = 2 sin(w s d / 2) / d, and the leapfrog step as Omega = 2 sin(w dt / 2) / dt (lightsmith's exp(-i w t)).

In a homogeneous host the background is the incident wave alone, and the discrete wave has a closed form: every
component's complex amplitude, referred to its own staggered position, obeys the continuum relations with k -> K,
w -> Omega:

    Omega^2 eps = Kx^2 + Ky^2 + Kz^2,   K . E = 0,   H = K x E / Omega,   Kz = 2 sin(kz dz / 2) / dz,

so F_c(z) = A_c exp(i kz (z_c - z_end)) with the tangential E at the incidence-side end node equal to
ehat_t. Over a substrate (the half-space z < z_int of another medium, the 1D axis then being z) the background
holds the reflected and transmitted waves too: per frequency the discrete equations along z, the Yee H and E
half-steps u_{k+1} = M_k u_k for u_k = (E_x, E_y at node k; H_x, H_y at k - 1/2), with every node's 1/eps from the
layer meshing (the substrate fraction f of the node's cell: 1/<eps> for tangential, <1/eps> for normal components,
eps the D-form ADE response ``discrete_eps``), are solved as one sparse linear system over the table rows, closed at
both ends by the end media's own discrete plane waves (the closed form above, two polarisations each way): the
incident one at the incidence-side end, outgoing ones (propagating away or decaying) at both. One global solve per
frequency has no growing exponentials to tame, so total internal reflection needs nothing special. A pulse g(t) (Gaussian envelope times sin(omega_c t), high-passed below omega_min / 3) is transformed,
multiplied by each component's response and transformed back; tables are kept every m steps (w h <= 0.2 at
omega_c + 3 omega_half), padded with zeros, and read per correction point with quintic interpolation at its own
delay. The time axis, pulse timing and padding follow layered.py:331-405 exactly, so the same box gets the same
source as the reference's.
"""

from __future__ import annotations

import numpy as np

from .media import Medium, discrete_eps
from .sources import TableSource, tfsf_slabs

OMEGA_H = 0.2  # table sample spacing: omega h at omega_c + 3 omega_half (layered.py:166)
STENCIL_LO, STENCIL_N = -2, 6  # quintic stencil -2 .. 3 (layered.py:164)
NAMES = ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz")


class ObliqueLayeredPlaneWave(TableSource):
    """TFSF box [lo, hi] (node-index units) fed with the grid's exact discrete plane wave along khat, polarised
    along ehat, in the host medium, or over the substrate ``sub`` filling z < z_int (layered.py:254-405 constructor
    contract; graded grids, mirror planes and beam foci are not supported here and raise). The incidence medium
    (the host for a wave going down, the substrate for one going up) must be non-dispersive.

    After construction: ``t_passed`` (the incident wave has passed every point, tables below 1e-11 of their
    peak), ``t_off`` (below 1e-6), ``n_inc`` (the host index), ``m`` (table stride in steps), ``view`` (the
    incident wave alone at the reference point, E_x, E_y, E_z, for a DFTMonitor's ``source``), ``background``
    and ``incident_reference`` (the frequency-domain background for flux bookkeeping)."""

    def __init__(self, *, shape, origin, spacing, dt, lo, hi, khat, ehat, inv_eps, host: Medium,
                 z_int: float = -np.inf, sub: Medium | None = None, omega_c, omega_half, max_steps: int,
                 omega_min: float | None = None, focus=None, t_focus: float = 0.0, mirror=(0, 0, 0), grid=None):
        if sub is not None and not np.isfinite(z_int):
            raise ValueError("a substrate needs a finite interface height z_int")
        if grid is not None:
            raise ValueError("graded grids are not supported (grid='uniform')")
        if any(mirror):
            raise ValueError("mirror planes are not supported (symmetry='off')")
        if focus is not None:
            raise ValueError("beam foci are not supported (plane waves only)")
        khat = np.asarray(khat, dtype=float) / np.linalg.norm(khat)
        ehat = np.asarray(ehat, dtype=float) / np.linalg.norm(ehat)
        a = 2 if sub is not None else int(np.argmax(np.abs(khat)))  # a substrate's interface is a z plane
        self.perm = perm = np.array([(j + a + 1) % 3 for j in range(3)])  # relabelled axis j is grid axis perm[j]
        self.inv = inv = np.argsort(perm)
        self.shape, self.dt = tuple(int(v) for v in shape), float(dt)
        sp_grid = np.broadcast_to(np.asarray(spacing, dtype=float), (3,)).copy()
        self.spacing = sp = sp_grid[perm]
        self.origin = np.asarray(origin, dtype=float)[perm]
        self.host, self.sub, self.z_int = host, sub, float(z_int)
        lo_g, hi_g = np.asarray(lo), np.asarray(hi)
        lo_r, hi_r = lo_g[perm], hi_g[perm]
        self.khat, self.ehat = khat[perm], ehat[perm]
        self.down = bool(self.khat[2] < 0)
        inc_m = host if (sub is None or self.down) else sub
        if inc_m.dispersive:
            raise ValueError("oblique incidence needs a non-dispersive incidence medium")
        self.eps = float(inc_m.eps_inf)
        self.n_inc = float(np.sqrt(self.eps))
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
        n_max = max(self.n_inc, np.sqrt(max(host.eps_inf, 1.0)), np.sqrt(max(sub.eps_inf, 1.0)) if sub else 0.0)
        span = self.tau_lead + 2 * self.t0 + 4 * nk * dz * n_max + 4 * self.tau_p
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
    def _kt(self, w):
        """Transverse and temporal grid wavenumbers Kx, Ky, Omega (n_w,)."""
        w = np.asarray(w, dtype=float)
        sp = self.spacing
        return (2 * np.sin(w * self.s[0] * sp[0] / 2) / sp[0], 2 * np.sin(w * self.s[1] * sp[1] / 2) / sp[1],
                2 * np.sin(w * self.dt / 2) / self.dt)

    def _wave(self, w, eps=None, down=None, et=None):
        """Component amplitudes (n_w, 6) (E_x, E_y, E_z, H_x, H_y, H_z in the relabelled frame, each referred to
        its own staggered position) and kz (n_w,) of the discrete plane wave in a uniform medium of permittivity
        eps (scalar or (n_w,), default the incidence medium's) going down (-z) or up, whose tangential E is et
        (default ehat_t). Going up means propagating with Re kz > 0 or, evanescent, decaying towards +z."""
        w = np.asarray(w, dtype=float)
        dz = self.spacing[2]
        Kx, Ky, Om = self._kt(w)
        eps = self.eps if eps is None else eps
        down = self.down if down is None else down
        et = self.ehat[:2] if et is None else et
        Kz = np.sqrt((Om ** 2 * eps - Kx ** 2 - Ky ** 2).astype(complex))
        if down:
            Kz = -Kz
        kz = 2 / dz * np.arcsin(Kz * dz / 2)
        ex = np.full(w.shape, et[0], dtype=complex)
        ey = np.full(w.shape, et[1], dtype=complex)
        ez = -(Kx * ex + Ky * ey) / Kz
        hx = (Ky * ez - Kz * ey) / Om
        hy = (Kz * ex - Kx * ez) / Om
        hz = (Kx * ey - Ky * ex) / Om
        return np.stack([ex, ey, ez, hx, hy, hz], axis=1), kz

    def _response(self, w) -> dict:
        """The discrete wave on the table rows: (n_w, nk) per component (E_x, E_y, H_z at z_k; E_z, H_x, H_y at
        z_k + dz / 2), and the incident wave at the end node, ``inc`` (n_w, 3): E_x, E_y (node), E_z (node + 1/2)."""
        if self.sub is not None:
            return self._layered(w)
        A, kz = self._wave(w)
        dz = self.spacing[2]
        z_node = self.z_nodes - self.z_end
        out = {}
        for i, nm in enumerate(NAMES):
            half = nm in ("Ez", "Hx", "Hy")
            out[nm] = A[:, i:i + 1] * np.exp(1j * kz[:, None] * (z_node + (dz / 2 if half else 0.0))[None, :])
        out["inc"] = np.stack([A[:, 0], A[:, 1], A[:, 2] * np.exp(0.5j * kz * dz)], axis=1)
        return out

    def layer_inv_eps(self, w, z, normal: bool) -> np.ndarray:
        """1/eps (n_w, n_z) of E nodes at heights z (relabelled frame) under the layer meshing: the substrate fraction
        f of the node's cell [z - dz/2, z + dz/2), <1/eps> for the normal component, 1/<eps> for tangential ones,
        eps the D-form ADE response of each medium (setup.py:187-234 meshing, layered.py:34-50)."""
        dz = self.spacing[2]
        f = np.clip((self.z_int - (np.asarray(z, dtype=float) - dz / 2)) / dz, 0.0, 1.0)[None, :]
        e_h = discrete_eps(self.host, w, self.dt)[:, None]
        e_s = discrete_eps(self.sub, w, self.dt)[:, None]
        return f / e_s + (1 - f) / e_h if normal else 1 / (f * e_s + (1 - f) * e_h)

    def _mode_u(self, w, eps, down: bool):
        """u = (E_x, E_y at a node; H_x, H_y half a cell below it) of the two discrete plane waves (tangential E
        (1, 0) and (0, 1)) of a uniform medium going down or up: (n_w, 4, 2), with their amplitudes (n_w, 6, 2)
        and kz (n_w, 2)."""
        dz = self.spacing[2]
        U = np.zeros((len(w), 4, 2), dtype=complex)
        A2 = np.zeros((len(w), 6, 2), dtype=complex)
        kz2 = np.zeros((len(w), 2), dtype=complex)
        for p, et in enumerate(((1.0, 0.0), (0.0, 1.0))):
            A, kz = self._wave(w, eps, down, et)
            back = np.exp(-0.5j * kz * dz)  # H_x, H_y are referred to the node + dz/2; u holds them at node - dz/2
            U[:, :, p] = np.stack([A[:, 0], A[:, 1], A[:, 3] * back, A[:, 4] * back], axis=1)
            A2[:, :, p], kz2[:, p] = A, kz
        return U, A2, kz2

    def _layered(self, w) -> dict:
        """The background over the substrate on the table rows (the dict of ``_response``): per frequency one sparse
        solve of the discrete equations along z (module docstring)."""
        from scipy.sparse import csr_matrix
        from scipy.sparse.linalg import spsolve

        w = np.asarray(w, dtype=float)
        nw, nk, dz = len(w), self.nk, self.spacing[2]
        Kx, Ky, Om = self._kt(w)
        eta_t = self.layer_inv_eps(w, self.z_nodes, False)            # E_x, E_y at z_k
        eta_z = self.layer_inv_eps(w, self.z_nodes + dz / 2, True)    # E_z at z_k + dz/2
        e_bot = 1 / eta_t[:, 0]   # the media continue beyond the table ends
        e_top = 1 / eta_t[:, -1]
        out_b, _, _ = self._mode_u(w, e_bot, True)     # outgoing below: going down
        out_t, _, _ = self._mode_u(w, e_top, False)    # outgoing above: going up
        A_inc, kz_inc = self._wave(w)                   # the incident wave in the incidence medium
        back = np.exp(-0.5j * kz_inc * dz)
        u_inc = np.stack([A_inc[:, 0], A_inc[:, 1], A_inc[:, 3] * back, A_inc[:, 4] * back], axis=1)
        # M_k (n_w, nk, 4, 4): the H half-step across node k, then the E half-step across k + 1/2
        Kx_, Ky_, Om_ = Kx[:, None], Ky[:, None], Om[:, None]
        MH = np.zeros((nw, nk, 4, 4), dtype=complex)
        MH[..., [0, 1, 2, 3], [0, 1, 2, 3]] = 1
        # H_z = (Kx E_y - Ky E_x) / Om;  H_x += dz (-i Om E_y / eta_t + i Kx H_z);  H_y += dz (i Ky H_z + i Om E_x / eta_t)
        MH[..., 2, 0] = dz * (-1j * Kx_ * Ky_ / Om_)
        MH[..., 2, 1] = dz * (-1j * Om_ / eta_t + 1j * Kx_ * Kx_ / Om_)
        MH[..., 3, 0] = dz * (1j * Om_ / eta_t - 1j * Ky_ * Ky_ / Om_)
        MH[..., 3, 1] = dz * (1j * Ky_ * Kx_ / Om_)
        ME = np.zeros((nw, nk, 4, 4), dtype=complex)
        ME[..., [0, 1, 2, 3], [0, 1, 2, 3]] = 1
        # E_z = -eta_z (Kx H_y - Ky H_x) / Om;  E_x += dz (i Om H_y + i Kx E_z);  E_y += dz (-i Om H_x + i Ky E_z)
        ME[..., 0, 2] = dz * (1j * Kx_ * eta_z * Ky_ / Om_)
        ME[..., 0, 3] = dz * (1j * Om_ - 1j * Kx_ * eta_z * Kx_ / Om_)
        ME[..., 1, 2] = dz * (-1j * Om_ + 1j * Ky_ * eta_z * Ky_ / Om_)
        ME[..., 1, 3] = dz * (-1j * Ky_ * eta_z * Kx_ / Om_)
        M = ME @ MH
        # unknowns: r_b (2), u_0 .. u_{nk-1} (4 each), r_t (2); rows: bottom closure (4), u_{k+1} - M_k u_k = 0
        # (4 (nk - 1)), top closure (4)
        n = 4 * nk + 4
        ui = lambda k: 2 + 4 * k  # noqa: E731  (first unknown of u_k)
        r_blk, c_blk = np.meshgrid(np.arange(4), np.arange(4), indexing="ij")
        rows, cols = [], []
        rows.append(np.repeat(np.arange(4), 1)); cols.append(ui(0) + np.arange(4))               # u_0
        rows.append(np.repeat(np.arange(4), 2)); cols.append(np.tile(np.arange(2), 4))           # - out_b r_b
        for k in range(nk - 1):
            r0 = 4 + 4 * k
            rows.append(r0 + np.arange(4)); cols.append(ui(k + 1) + np.arange(4))                # u_{k+1}
            rows.append(r0 + r_blk.ravel()); cols.append(ui(k) + c_blk.ravel())                  # - M_k u_k
        rt = 4 * nk
        rows.append(rt + np.arange(4)); cols.append(ui(nk - 1) + np.arange(4))                   # u_{nk-1}
        rows.append(rt + np.repeat(np.arange(4), 2)); cols.append(n - 2 + np.tile(np.arange(2), 4))  # - out_t r_t
        rows, cols = np.concatenate(rows), np.concatenate(cols)
        U = np.zeros((nw, nk, 4), dtype=complex)
        ones = np.ones((nk - 1, 4))
        for q in range(nw):
            mid = np.concatenate([ones, -M[q, :nk - 1].reshape(nk - 1, 16)], axis=1).ravel()
            vals = np.concatenate([np.ones(4), -out_b[q].ravel(), mid, np.ones(4), -out_t[q].ravel()])
            A = csr_matrix((vals, (rows, cols)), shape=(n, n))
            rhs = np.zeros(n, dtype=complex)
            if self.down:
                rhs[rt:rt + 4] = u_inc[q]
            else:
                rhs[0:4] = u_inc[q]
            x = spsolve(A.tocsc(), rhs)
            U[q] = x[2:2 + 4 * nk].reshape(nk, 4)
        Ex, Ey = U[..., 0], U[..., 1]
        # H_x, H_y at z_k + dz/2: the H half-step across node k (u_{k+1}'s H part)
        Hx = U[..., 2] + MH[..., 2, 0] * Ex + MH[..., 2, 1] * Ey
        Hy = U[..., 3] + MH[..., 3, 0] * Ex + MH[..., 3, 1] * Ey
        Hz = (Kx_ * Ey - Ky_ * Ex) / Om_
        Ez = -eta_z * (Kx_ * Hy - Ky_ * Hx) / Om_
        inc = np.stack([A_inc[:, 0], A_inc[:, 1], A_inc[:, 2] * np.exp(0.5j * kz_inc * dz)], axis=1)
        return {"Ex": Ex, "Ey": Ey, "Ez": Ez, "Hx": Hx, "Hy": Hy, "Hz": Hz, "inc": inc}

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
        if self.sub is not None:  # over a substrate: the solved rows (NaN outside the table's z range)
            w = np.asarray(omega, dtype=float)
            f = self._layered(w)
            rows = np.floor(pos[:, 2] + 1e-9).astype(np.int64) - self.k0
            ok = (rows >= 0) & (rows < self.nk)
            phase = np.exp(1j * w[:, None] * ((self._par(pos[:, :2]) - self.r_c) @ self.s)[None, :])
            vals = np.zeros((len(w), len(comp)), dtype=complex)
            for c in range(3):
                sel = rc == c
                vals[:, sel] = f[kind + "xyz"[c]][:, np.clip(rows[sel], 0, self.nk - 1)]
            return np.where(ok[None, :], vals * phase, np.nan)
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
