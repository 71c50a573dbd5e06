"""FP64 restatement of the conformal vertex tensor (oracle; test infrastructure only).

``TensorMesh`` and ``TensorGroup`` follow engine.py:140-181 (the data) and engine.py:796-923 (_Harmonic,
_Arithmetic, _Tensor: the per-step D-form update of interface nodes with the smoothed 1/eps tensor applied at grid
vertices). ``vertex_tensor`` follows setup.py:97, 293-420 (the tensor vertices of interfaces with a dispersive
medium and their node/slot tables) for a uniform grid without mirror planes or a layered background; it generates
test inputs. Arrays use the reference's flat indexing (engine.py:103-105).
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .fdtd import flat_index
from .media import pole_coefficients

_TEST = np.stack(np.meshgrid(*[np.array([-0.5, 0.0, 0.5])] * 3, indexing="ij"), axis=-1).reshape(-1, 3)


@dataclass
class TensorMesh:  # engine.py:140-181
    media: list
    frac: np.ndarray  # (S, nV)
    normal: np.ndarray  # (nV, 3)
    node_comp: np.ndarray  # (nN,)
    node_idx3: np.ndarray  # (nN, 3)
    vert_nodes: np.ndarray  # (nV, 3, 2)
    shift: np.ndarray  # (nV, 3, 2, 3)
    slot_vert: np.ndarray  # (nN, 2)
    slot_medium: np.ndarray  # (nN, 2)
    shape: tuple
    sign: np.ndarray | None = None
    vert_idx3: np.ndarray | None = None

    @property
    def idx(self) -> np.ndarray:
        i = self.node_idx3
        return np.asarray(flat_index(self.node_comp, i[:, 0], i[:, 1], i[:, 2], self.shape), dtype=np.int64)


def _canonical(x, shape, periodic):
    """Real indices and period shifts of possibly wrapped indices (setup.py:293-303)."""
    x = np.array(x, dtype=np.int64)
    shift = np.zeros_like(x)
    for a in range(3):
        if periodic[a]:
            P = shape[a] - 2
            c = 1 + np.mod(x[:, a] - 1, P)
            shift[:, a] = (x[:, a] - c) // P
            x[:, a] = c
    return x, shift


def _pixel_average(slot_fn, pos, dx, S: int, samples: int):
    """Medium fractions (S, n) and unit normals (n, 3) of cubes of side dx centred on pos (setup.py:306-322): the
    normal is the first moment of the dominant medium's indicator over the sub-samples."""
    sub = (np.arange(samples) + 0.5) / samples - 0.5
    sub3 = np.stack(np.meshgrid(sub, sub, sub, indexing="ij"), axis=-1).reshape(-1, 3)
    sl = slot_fn(pos[:, None, :] + sub3[None, :, :] * dx)
    frac = np.stack([(sl == s).mean(axis=1) for s in range(S)])
    dom = np.argmax(frac, axis=0)
    chi = (sl == dom[:, None]).astype(float)
    mom = (chi - chi.mean(axis=1, keepdims=True)) @ sub3 / dx
    nrm = np.linalg.norm(mom, axis=1)
    nhat = np.where(nrm[:, None] > 1e-12, mom / np.where(nrm > 0, nrm, 1)[:, None], 0.0)
    return frac, nhat


def vertex_tensor(slot_fn, media, shape, origin, dx: float, roi_lo, roi_hi, samples: int = 8,
                  periodic=(False, False, False), background=()):
    """The TensorMesh of interfaces with a dispersive medium, or None (setup.py:325-420, uniform grid). Vertices
    whose cube holds only the ``background`` media are not mixed (setup.py:359-363)."""
    shape = tuple(shape)
    S = len(media)
    disp = np.array([m.dispersive for m in media])
    dx3 = np.full(3, float(dx))
    rng = [np.arange(1, n - 1) if periodic[a] else np.arange(max(int(roi_lo[a]), 1), min(int(roi_hi[a]), n - 2) + 1)
           for a, n in enumerate(shape)]
    V = np.stack(np.meshgrid(*rng, indexing="ij"), axis=-1).reshape(-1, 3)
    pos = origin + V * dx3
    center = slot_fn(pos)
    mixed = np.zeros(len(V), dtype=bool)
    for o in _TEST:
        if np.any(o):
            mixed |= slot_fn(pos + o * dx3) != center
    Vm = V[mixed]
    if not len(Vm):
        return None
    frac, nhat = _pixel_average(slot_fn, pos[mixed], dx3, S, samples)
    if len(background):
        part = np.delete(frac, list(background), axis=0).sum(axis=0) > 0
        Vm, frac, nhat = Vm[part], frac[:, part], nhat[part]
        if not len(Vm):
            return None
    seed = (frac[disp] > 0).any(axis=0) & (frac.max(axis=0) < 1)
    if not seed.any():
        return None

    def key(x):
        return (x[:, 0] * shape[1] + x[:, 1]) * shape[2] + x[:, 2]

    kv = key(Vm)
    order = np.argsort(kv)
    ks = kv[order]

    def lookup(x):
        k = key(x)
        j = np.clip(np.searchsorted(ks, k), 0, len(ks) - 1)
        return np.where(ks[j] == k, order[j], -1)

    eye = np.eye(3, dtype=np.int64)
    inT = seed.copy()
    front = np.nonzero(seed)[0]
    while len(front):
        nb = np.concatenate([_canonical(Vm[front] + s * eye[a], shape, periodic)[0] for a in range(3) for s in (-1, 1)])
        j = lookup(nb)
        j = np.unique(j[j >= 0])
        front = j[~inT[j]]
        inT[front] = True
    T = Vm[inT]
    fracT, nT = frac[:, inT], nhat[inT]
    nV = len(T)
    raw = np.stack([np.stack([T - eye[d], T], axis=1) for d in range(3)], axis=1)  # (nV, 3, 2, 3)
    comp = np.broadcast_to(np.arange(3)[None, :, None], (nV, 3, 2)).reshape(-1)
    real, shift = _canonical(raw.reshape(-1, 3), shape, periodic)
    nkey = comp * int(np.prod(shape)) + key(real)
    _, first, inv = np.unique(nkey, return_index=True, return_inverse=True)
    node_comp, node_idx3 = comp[first], real[first]
    nN = len(first)
    tkey = key(T)
    torder = np.argsort(tkey)
    tks = tkey[torder]
    slot_vert = np.full((nN, 2), -1, dtype=np.int64)
    slot_medium = np.zeros((nN, 2), dtype=np.int64)
    for k in range(2):
        vx, _ = _canonical(node_idx3 + k * eye[node_comp], shape, periodic)
        kk = key(vx)
        j = np.clip(np.searchsorted(tks, kk), 0, len(tks) - 1)
        hit = tks[j] == kk
        slot_vert[hit, k] = torder[j[hit]]
        if (~hit).any():
            slot_medium[~hit, k] = slot_fn(origin + vx[~hit] * dx3)
    return TensorMesh(media=list(media), frac=fracT, normal=nT, node_comp=node_comp, node_idx3=node_idx3,
                      vert_nodes=inv.reshape(nV, 3, 2), shift=shift.reshape(nV, 3, 2, 3), slot_vert=slot_vert,
                      slot_medium=slot_medium, shape=shape, vert_idx3=T)


def _advance(pairs, coefs, e):
    """P^{n+1} = c1 P^n + c2 P^{n-1} + c3 e per pole; returns sum P^{n+1} (engine.py:760-767)."""
    tot = None
    for (P, Pp), (c1, c2, c3) in zip(pairs, coefs):
        new = c1 * P + c2 * Pp + c3 * e
        Pp[...] = P
        P[...] = new
        tot = P.copy() if tot is None else tot + P
    return tot


class Harmonic:
    """y = <1/eps>[u], series combination, each medium's poles driven by its own E_s (engine.py:796-816)."""

    def __init__(self, eng, media, F):
        n = F.shape[1]
        self.h0 = sum(F[s] / m.eps_inf for s, m in enumerate(media) if not m.dispersive) + 0 * F[0]
        self.disp = [s for s, m in enumerate(media) if m.dispersive and np.any(F[s] > 0)]
        self.f = {s: F[s] for s in self.disp}
        self.inv = {s: 1.0 / media[s].eps_inf for s in self.disp}
        self.coef = {s: [pole_coefficients(p, eng.dt) for p in media[s].poles] for s in self.disp}
        self.E = {s: np.zeros(n, eng.fdt) for s in self.disp}
        self.P = {s: [(np.zeros(n, eng.fdt), np.zeros(n, eng.fdt)) for _ in media[s].poles] for s in self.disp}

    def __call__(self, u):
        y = u * self.h0
        for s in self.disp:
            tot = _advance(self.P[s], self.coef[s], self.E[s])
            self.E[s][...] = (u - tot) * self.inv[s]
            y = y + self.f[s] * self.E[s]
        return y


class Arithmetic:
    """y = (1/<eps>)[u], parallel combination, all poles driven by y (engine.py:819-836)."""

    def __init__(self, eng, media, F):
        n = F.shape[1]
        self.disp = [s for s, m in enumerate(media) if m.dispersive and np.any(F[s] > 0)]
        self.f = {s: F[s] for s in self.disp}
        self.inv_a = 1.0 / sum(F[s] * m.eps_inf for s, m in enumerate(media))
        self.coef = {s: [pole_coefficients(p, eng.dt) for p in media[s].poles] for s in self.disp}
        self.y = np.zeros(n, eng.fdt)
        self.P = {s: [(np.zeros(n, eng.fdt), np.zeros(n, eng.fdt)) for _ in media[s].poles] for s in self.disp}

    def __call__(self, u):
        acc = u
        for s in self.disp:
            acc = acc - self.f[s] * _advance(self.P[s], self.coef[s], self.y)
        self.y[...] = acc * self.inv_a
        return self.y


class TensorGroup:
    """D-form update of a TensorMesh's nodes (engine.py:839-923), uniform grid without mirror planes: per step
    D += E_now - E_prev; per vertex the mean D of each component pair, s = n . D~; per tensor slot
    sigma = s + n_c (D_p - D~_c) and (1/2)(b[D_p - n_c sigma] + n_c a[sigma]) (harmonic a, arithmetic b, each slot
    with its own filter state); plain slots (1/2) medium[D_p]; E_p = the sum of its two slots."""

    def __init__(self, eng, g: TensorMesh):
        if g.sign is not None:
            raise ValueError("mirror planes are not restated")
        nN, nV = len(g.node_comp), len(g.normal)
        self.idx = g.idx
        self.D = np.zeros(nN, eng.fdt)
        self.E_prev = np.zeros(nN, eng.fdt)
        self.out = np.zeros(2 * nN, eng.fdt)
        L = np.array([(n - 2) * d for n, d in zip(eng.shape, eng.spacing)])
        ph = np.exp(1j * (g.shift * (np.asarray(eng.bloch) * L)).sum(-1))  # (nV, 3, 2)
        self.complex = eng.complex
        self.vn = g.vert_nodes.reshape(-1)
        self.nV = nV
        self.ph = ph.reshape(-1) if self.complex else None
        self.n = g.normal.astype(float)
        sv = g.slot_vert
        tk, tnode = np.nonzero((sv >= 0).T)
        tv = sv[tnode, tk]
        tc = g.node_comp[tnode]
        side = 1 - tk
        self.t_pos = tk * nN + tnode
        self.t_v = tv
        self.t_vc = tv * 3 + tc
        self.t_g = (tv * 3 + tc) * 2 + side
        self.t_nc = self.n[tv, tc]
        self.t_conj = np.conj(ph[tv, tc, side]) if self.complex else None
        F = g.frac[:, tv]
        self.a = Harmonic(eng, g.media, F)
        self.b = Arithmetic(eng, g.media, F)
        pk, pnode = np.nonzero((sv < 0).T)
        self.p_pos = pk * nN + pnode
        self.p_node = pnode
        onehot = np.zeros((len(g.media), len(pnode)))
        onehot[g.slot_medium[pnode, pk], np.arange(len(pnode))] = 1.0
        self.plain = Harmonic(eng, g.media, onehot)
        self.has_plain = len(pnode) > 0
        self.nN = nN

    def step(self, eng):
        e_now = eng.E_flat[self.idx]
        self.D += e_now - self.E_prev
        dv = self.D[self.vn]
        if self.complex:
            dv = dv * self.ph
        dv = dv.reshape(self.nV, 3, 2)
        dbar = 0.5 * (dv[..., 0] + dv[..., 1])
        s = np.sum(self.n * dbar, axis=1)
        dp = dv.reshape(-1)[self.t_g]
        nc = self.t_nc
        sigma = s[self.t_v] + nc * (dp - dbar.reshape(-1)[self.t_vc])
        contrib = 0.5 * (self.b(dp - nc * sigma) + nc * self.a(sigma))
        if self.complex:
            contrib = contrib * self.t_conj
        self.out[self.t_pos] = contrib
        if self.has_plain:
            self.out[self.p_pos] = 0.5 * self.plain(self.D[self.p_node])
        e_new = self.out[:self.nN] + self.out[self.nN:]
        eng.E_flat[self.idx] = e_new
        self.E_prev[...] = e_new
