"""Material rasteriser: particles -> the engine's material arrays, with subpixel smoothing at interfaces
(setup.py:110-273 of the reference) and the conformal vertex tensor at interfaces with a dispersive medium
(setup.py:325-420).

The point classification -- which medium holds a point -- is the expensive part (every interface cell is
supersampled samples^3 times); it runs as torch tensor code on the engine's device, so on the MI355X the
half-cell lattice and the supersamples of a 320^3 scene are classified on the GPU. Index bookkeeping (the
interface node lists, the vertex tensor's node and slot tables) is host NumPy over surface-sized arrays.

Modes (``FDTDOptions.mesh_refinement``):
  staircase       the medium at each E position;
  volume_average  interface cells get <eps> over the cell;
  subpixel        the diagonal of the Kottke tensor, 1/eps_cc = w <1/eps> + (1 - w) / <eps>, w = n_c^2, n the
                  first moment of the dominant medium's indicator over the cell (an interface normal);
  conformal       subpixel plus the off-diagonal terms n_c n_d (<1/eps> - 1/<eps>) of non-dispersive cells, and
                  the full tensor at grid vertices next to a dispersive medium (TensorMesh).
A non-dispersive pure cell sets 1/eps; a dispersive pure cell becomes a D-form node (1/eps = 1) of its medium's
pure group; an interface cell holding a dispersive medium becomes a node of the mixed group (two-path ADE).
"""

from __future__ import annotations

import numpy as np
import torch

from .engine import DispersiveGroup, OffDiagonal, TensorMesh, flat_index
from .media import Medium

MESH_MODES = ("staircase", "volume_average", "subpixel", "conformal")
# the 26 neighbours at half-cell offsets: a cell is an interface cell when any of them holds another medium
_OFFSETS = [(a, b, c) for a in (-1, 0, 1) for b in (-1, 0, 1) for c in (-1, 0, 1) if (a, b, c) != (0, 0, 0)]


# --- point classification ---------------------------------------------------------------------------------------

class ParticleSlots:
    """pts (..., 3) world nm (torch float64) -> medium index (int64), 0 the host. The first particle listed wins
    where particles overlap (setup.py:707-716). Shapes: sphere, spheroid (axis z), cylinder (axis z), layered
    sphere (core and shells); their own region() is used for anything else."""

    def __init__(self, particles, slot_of):
        self.parts = []
        for p, slots in zip(particles, slot_of):
            self.parts.append((p.shape, np.asarray(p.position_nm, dtype=float), list(slots)))

    @staticmethod
    def _region(shape, q: torch.Tensor) -> torch.Tensor:
        kind = getattr(shape, "kind", None)
        r2 = (q * q).sum(-1)
        if kind == "sphere":
            return torch.where(r2 <= shape.radius_nm ** 2, 0, -1)
        if kind == "spheroid":
            a, c = shape.semi_axis_xy_nm, shape.semi_axis_z_nm
            v = (q[..., 0] ** 2 + q[..., 1] ** 2) / a ** 2 + q[..., 2] ** 2 / c ** 2
            return torch.where(v <= 1, 0, -1)
        if kind == "cylinder":
            ok = (q[..., 0] ** 2 + q[..., 1] ** 2 <= shape.radius_nm ** 2) & (q[..., 2].abs() <= shape.length_nm / 2)
            return torch.where(ok, 0, -1)
        if kind == "layered_sphere":
            radii = torch.as_tensor(np.asarray(shape.radii_nm, dtype=float), device=q.device)
            r = torch.sqrt(r2)
            out = torch.searchsorted(radii, r.contiguous(), right=False)
            return torch.where(r <= radii[-1], out, -1)
        reg = np.asarray(shape.region(q.detach().cpu().numpy()))
        return torch.as_tensor(reg, device=q.device)

    def __call__(self, pts: torch.Tensor) -> torch.Tensor:
        out = torch.zeros(pts.shape[:-1], dtype=torch.int64, device=pts.device)
        for shape, pos, slots in self.parts:
            reg = self._region(shape, pts - torch.as_tensor(pos, device=pts.device))
            for r, s in enumerate(slots):
                out = torch.where((reg == r) & (out == 0), s, out)
        return out


class OnSubstrate:
    """Points -> medium slot of particles over a substrate: where ``inner`` finds the host (slot 0) below the plane
    z = z_int, the substrate's slot (scattering.py:83-90)."""

    def __init__(self, inner, z_int: float, sub_slot: int):
        self.inner, self.z_int, self.sub_slot = inner, float(z_int), int(sub_slot)

    def __call__(self, pts: torch.Tensor) -> torch.Tensor:
        out = self.inner(pts)
        return torch.where((out == 0) & (pts[..., 2] < self.z_int), self.sub_slot, out)


def _classify(slot_fn, pts: torch.Tensor, chunk: int = 1 << 22) -> torch.Tensor:
    """slot_fn over a (n, 3) point list in chunks."""
    out = torch.empty(pts.shape[0], dtype=torch.int64, device=pts.device)
    for a in range(0, pts.shape[0], chunk):
        out[a:a + chunk] = slot_fn(pts[a:a + chunk])
    return out


def _sub3(samples: int, dev) -> torch.Tensor:
    sub = (torch.arange(samples, dtype=torch.float64, device=dev) + 0.5) / samples - 0.5
    return torch.stack(torch.meshgrid(sub, sub, sub, indexing="ij"), dim=-1).reshape(-1, 3)


def _pixel_stats(slot_fn, ctr: torch.Tensor, size: torch.Tensor, S: int, samples: int, chunk: int = 4096):
    """Medium fractions (S, n) and unit normals (n, 3) (0 where undefined) of boxes of side `size` (3,) centred on
    ctr (n, 3): each box sampled on a samples^3 lattice; the normal is the first moment of the dominant medium's
    indicator (pointing out of it), normalised."""
    dev = ctr.device
    sub3 = _sub3(samples, dev)
    n = ctr.shape[0]
    frac = torch.zeros((S, n), dtype=torch.float64, device=dev)
    nhat = torch.zeros((n, 3), dtype=torch.float64, device=dev)
    med = torch.arange(S, device=dev)
    for a in range(0, n, chunk):
        c = ctr[a:a + chunk]
        sl = slot_fn(c[:, None, :] + sub3[None, :, :] * size)  # (m, samples^3)
        f = (sl[None, :, :] == med[:, None, None]).to(torch.float64).mean(-1)  # (S, m)
        dom = torch.argmax(f, dim=0)
        chi = (sl == dom[:, None]).to(torch.float64)
        mom = (chi - chi.mean(1, keepdim=True)) @ sub3 / size
        nrm = torch.linalg.norm(mom, dim=1)
        frac[:, a:a + chunk] = f
        nhat[a:a + chunk] = torch.where(nrm[:, None] > 1e-12, mom / torch.where(nrm > 0, nrm, 1.0)[:, None], 0.0)
    return frac, nhat


# --- planar background ------------------------------------------------------------------------------------------

def planar_background(shape, origin, dx, z_int: float, host: Medium, sub: Medium) -> np.ndarray:
    """inv_eps [3, *shape] of the host above the plane z = z_int and a non-dispersive substrate below it, meshed as
    the layered source tables assume (substrate.py:25-70 for non-dispersive media): with f the substrate's share of
    an E node's cell along z, E_z takes <1/eps> = f/eps_s + (1-f)/eps_h and E_x, E_y take 1/<eps>."""
    if host.dispersive or sub.dispersive:
        raise ValueError("the planar background is meshed here for non-dispersive host and substrate only")
    shape = tuple(int(v) for v in shape)
    dz = float(np.broadcast_to(np.asarray(dx, dtype=float), (3,))[2])
    inv_eps = np.empty((3,) + shape)
    k = np.arange(shape[2], dtype=float)
    e_h, e_s = host.eps_inf, sub.eps_inf
    for c in range(3):
        z = origin[2] + (k + (0.5 if c == 2 else 0.0)) * dz  # E_c node heights
        f = np.clip((z_int - (z - dz / 2)) / dz, 0.0, 1.0)
        inv_eps[c] = (f / e_s + (1 - f) / e_h if c == 2 else 1 / (f * e_s + (1 - f) * e_h))[None, None, :]
    return inv_eps


# --- voxelisation -----------------------------------------------------------------------------------------------

def voxelise(slot_fn, media: list[Medium], shape, origin, dx, roi_lo, roi_hi, *, mode: str = "conformal",
             samples: int = 8, device="cpu", periodic=(False, False, False), base=None, background=()):
    """(inv_eps [3, *shape] float64 with 1 at D-form nodes, groups, stats, OffDiagonal or None) for the media
    slot_fn assigns (0 = the host, which must be non-dispersive); only nodes in [roi_lo, roi_hi) are examined
    (setup.py:110-273; uniform grid, no mirror planes). ``base``: the inv_eps of a planar background meshed
    beforehand (``planar_background``) and ``background`` the media slots it holds: a node whose cell holds
    background media only keeps its base value, so a planar interface is meshed alike inside and outside the
    examined region, and the vertex tensor does not grow along it."""
    if mode not in MESH_MODES:
        raise ValueError(f"mesh refinement {mode!r} is not one of {MESH_MODES}")
    if media[0].dispersive:
        raise ValueError("the host medium (slot 0) must be non-dispersive")
    dev = torch.device(device)
    shape = tuple(int(v) for v in shape)
    dx3 = np.broadcast_to(np.asarray(dx, dtype=float), (3,)).copy()
    origin = np.asarray(origin, dtype=float)
    roi_lo, roi_hi = np.asarray(roi_lo, dtype=np.int64), np.asarray(roi_hi, dtype=np.int64)
    S = len(media)
    disp = np.array([m.dispersive for m in media])
    eps_inf = np.array([m.eps_inf for m in media])
    inv_eps = np.full((3,) + shape, 1.0 / eps_inf[0]) if base is None else np.array(base, dtype=float)
    inv_flat = inv_eps.reshape(-1)
    is_bg = np.zeros(S, dtype=bool)
    is_bg[list(background)] = True
    n_roi = roi_hi - roi_lo
    # the half-cell lattice origin + (m / 2) dx, m = 2 lo - 1 .. 2 hi: every E position of the examined nodes
    # and every corner and edge midpoint of their cells, classified once
    ax = [torch.as_tensor(origin[a] + np.arange(2 * roi_lo[a] - 1, 2 * roi_hi[a] + 1) / 2 * dx3[a], device=dev)
          for a in range(3)]
    lat = torch.empty(tuple(len(v) for v in ax), dtype=torch.int64, device=dev)
    yz = torch.stack(torch.meshgrid(ax[1], ax[2], indexing="ij"), dim=-1).reshape(-1, 2)
    step = max(1, (1 << 22) // yz.shape[0])
    for i in range(0, len(ax[0]), step):
        x = ax[0][i:i + step]
        pts = torch.cat([x.repeat_interleave(yz.shape[0])[:, None], yz.repeat(len(x), 1)], dim=1)
        lat[i:i + len(x)] = slot_fn(pts).reshape(len(x), lat.shape[1], lat.shape[2])
    grids = np.meshgrid(*[np.arange(roi_lo[a], roi_hi[a]) for a in range(3)], indexing="ij")
    idx3 = np.stack(grids, axis=-1).reshape(-1, 3)
    subpixel = mode != "staircase"
    pure_idx = {s: [] for s in range(S) if disp[s]}
    mixed_idx, mixed_f, mixed_w = [], [], []
    od_idx3, od_comp, od_coef = [], [], []
    n_mixed = 0
    size = torch.as_tensor(dx3, device=dev)
    for c in range(3):
        def at(o, c=c):
            sl = tuple(slice(1 + (a == c) + o[a], 1 + (a == c) + o[a] + 2 * int(n_roi[a]), 2) for a in range(3))
            return lat[sl].reshape(-1)

        center = at((0, 0, 0))
        mixed = torch.zeros_like(center, dtype=torch.bool)
        if subpixel:
            for o in _OFFSETS:
                mixed |= at(o) != center
        center_h, mixed_h = center.cpu().numpy(), mixed.cpu().numpy()
        flat = flat_index(c, idx3[:, 0], idx3[:, 1], idx3[:, 2], shape)
        for s in range(1, S):
            sel = (~mixed_h) & (center_h == s)
            if is_bg[s] or not sel.any():
                continue
            if disp[s]:
                pure_idx[s].append(flat[sel])
                inv_flat[flat[sel]] = 1.0
            else:
                inv_flat[flat[sel]] = 1.0 / eps_inf[s]
        mi = np.nonzero(mixed_h)[0]
        if not len(mi):
            continue
        off = np.zeros(3)
        off[c] = 0.5
        ctr = torch.as_tensor(origin + (idx3[mi] + off) * dx3, device=dev)
        frac_t, nhat_t = _pixel_stats(slot_fn, ctr, size, S, samples)
        frac, nhat = frac_t.cpu().numpy(), nhat_t.cpu().numpy()
        if is_bg.any():  # cells of background media only keep the base
            part = frac[~is_bg].sum(axis=0) > 0
            mi, frac, nhat = mi[part], frac[:, part], nhat[part]
            if not len(mi):
                continue
        n_mixed += len(mi)
        has_n = np.linalg.norm(nhat, axis=1) > 0
        w = np.where(has_n, nhat[:, c] ** 2, 1 / 3)
        if mode == "volume_average":
            w = np.zeros_like(w)
        has_disp = (frac[disp] > 0).any(axis=0) if disp.any() else np.zeros(len(mi), bool)
        f_idx = flat[mi]
        nd = ~has_disp
        if nd.any():
            harm = (frac[:, nd] / eps_inf[:, None]).sum(axis=0)
            arit = (frac[:, nd] * eps_inf[:, None]).sum(axis=0)
            inv_flat[f_idx[nd]] = w[nd] * harm + (1 - w[nd]) / arit
            if mode == "conformal":
                coef = nhat[nd] * nhat[nd][:, c:c + 1] * (harm - 1 / arit)[:, None]
                coef[:, c] = 0
                keep = np.any(np.abs(coef) > 1e-9, axis=1)
                od_idx3.append(idx3[mi][nd][keep])
                od_comp.append(np.full(int(keep.sum()), c))
                od_coef.append(coef[keep])
        if has_disp.any():
            mixed_idx.append(f_idx[has_disp])
            mixed_f.append(frac[:, has_disp])
            mixed_w.append(w[has_disp])
            inv_flat[f_idx[has_disp]] = 1.0
    groups = []
    for s, parts in pure_idx.items():
        if parts:
            idx = np.concatenate(parts)
            groups.append(DispersiveGroup(idx, [media[s]], np.ones((1, len(idx))), None))
    if mixed_idx:
        f = np.concatenate(mixed_f, axis=1)
        used = [s for s in range(S) if np.any(f[s] > 0)]
        groups.append(DispersiveGroup(np.concatenate(mixed_idx), [media[s] for s in used], f[used],
                                      np.concatenate(mixed_w)))
    offdiag = None
    if od_comp and sum(len(v) for v in od_comp):
        offdiag = OffDiagonal(np.concatenate(od_idx3), np.concatenate(od_comp), np.concatenate(od_coef))
    tm = None
    if mode == "conformal" and disp.any():
        tm = vertex_tensor(slot_fn, media, shape, origin, dx3, roi_lo, roi_hi, samples=samples, device=dev,
                           periodic=periodic, background=background)
    if tm is not None:
        taken = tm.idx
        inv_flat[taken] = 1.0
        kept = []
        for g in groups:
            keep = ~np.isin(g.idx, taken)
            if keep.any():
                kept.append(DispersiveGroup(g.idx[keep], g.media, g.frac[:, keep], None if g.w is None else g.w[keep]))
        groups = kept + [tm]
        if offdiag is not None:
            i = offdiag.idx3
            keep = ~np.isin(flat_index(offdiag.comp, i[:, 0], i[:, 1], i[:, 2], shape), taken)
            offdiag = OffDiagonal(i[keep], offdiag.comp[keep], offdiag.coef[keep]) if keep.any() else None
    stats = {"interface_cells": int(n_mixed), "dispersive_cells": int(sum(len(g.idx) for g in groups)),
             "mesh_refinement": mode, "offdiagonal_cells": 0 if offdiag is None else len(offdiag.comp),
             "tensor_vertices": 0 if tm is None else len(tm.normal)}
    return inv_eps, groups, stats, offdiag


# --- conformal vertex tensor --------------------------------------------------------------------------------------

def _wrap(x: np.ndarray, shape, periodic) -> tuple[np.ndarray, np.ndarray]:
    """(real index, period count) of grid indices (n, 3) that may lie on or past a periodic axis' ghost planes:
    along a periodic axis the real indices are 1 .. n - 2, and index i is real index 1 + (i - 1) mod (n - 2)
    shifted by floor((i - 1) / (n - 2)) periods (setup.py:292-303)."""
    x = np.array(x, dtype=np.int64)
    shift = np.zeros_like(x)
    for a in range(3):
        if periodic[a]:
            shift[:, a], r = np.divmod(x[:, a] - 1, shape[a] - 2)
            x[:, a] = 1 + r
    return x, shift


def vertex_tensor(slot_fn, media: list[Medium], shape, origin, dx, roi_lo, roi_hi, *, samples: int = 8,
                  device="cpu", periodic=(False, False, False), background=()) -> TensorMesh | None:
    """The TensorMesh of interfaces with a dispersive medium, or None (setup.py:325-420; uniform grid, no mirror
    planes). Tensor vertices: every vertex whose cube (side dx, centred on it) holds an interface with a dispersive
    medium, grown over the mixed vertices connected to them through shared nodes, so that each node next to a
    tensor vertex has at its other end another tensor vertex or a single medium. Each node's two slots are its
    vertices at the node index (slot 0) and at index + e_c (slot 1); a slot without a tensor vertex takes the
    medium at that vertex. Along a periodic axis every real vertex is examined and indices past the cell wrap to
    their real node, the period count kept in ``shift`` (the engine applies the Bloch phase of it). A vertex whose
    cube holds ``background`` media only (a planar background meshed apart) is not a mixed vertex."""
    dev = torch.device(device)
    shape = tuple(int(v) for v in shape)
    S = len(media)
    disp = np.array([m.dispersive for m in media])
    dx3 = np.broadcast_to(np.asarray(dx, dtype=float), (3,)).copy()
    origin = np.asarray(origin, dtype=float)
    rng = [np.arange(1, n - 1) if periodic[a] else np.arange(max(int(roi_lo[a]), 1), min(int(roi_hi[a]), n - 2) + 1)
           for a, n in enumerate(shape)]
    V = np.stack(np.meshgrid(*rng, indexing="ij"), axis=-1).reshape(-1, 3)
    pos = torch.as_tensor(origin + V * dx3, device=dev)
    size = torch.as_tensor(dx3, device=dev)
    center = _classify(slot_fn, pos)
    mixed = torch.zeros_like(center, dtype=torch.bool)
    for o in _OFFSETS:
        mixed |= _classify(slot_fn, pos + 0.5 * torch.as_tensor(o, dtype=torch.float64, device=dev) * size) != center
    mixed = mixed.cpu().numpy()
    Vm = V[mixed]
    if not len(Vm):
        return None
    frac_t, nhat_t = _pixel_stats(slot_fn, pos[torch.as_tensor(mixed, device=dev)], size, S, samples)
    frac, nhat = frac_t.cpu().numpy(), nhat_t.cpu().numpy()
    if len(background):
        fg = np.ones(S, dtype=bool)
        fg[list(background)] = False
        part = frac[fg].sum(axis=0) > 0
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
    eye = np.eye(3, dtype=np.int64)
    # grow the seed over mixed vertices that share a node with it (breadth first)
    inT = seed.copy()
    front = np.nonzero(seed)[0]
    while len(front):
        nb = _wrap(np.concatenate([Vm[front] + s * eye[a] for a in range(3) for s in (-1, 1)]), shape, periodic)[0]
        k = key(nb)
        j = np.clip(np.searchsorted(ks, k), 0, len(ks) - 1)
        j = np.unique(order[j][ks[j] == k])
        front = j[~inT[j]]
        inT[front] = True
    T = Vm[inT]
    nV = len(T)
    # a vertex touches the E_d nodes at T - e_d (its minus side) and T (its plus side)
    raw = np.stack([np.stack([T - eye[d], T], axis=1) for d in range(3)], axis=1)  # (nV, 3, 2, 3)
    comp = np.broadcast_to(np.arange(3)[None, :, None], (nV, 3, 2)).reshape(-1)
    real, shift = _wrap(raw.reshape(-1, 3), shape, periodic)
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
        vx = _wrap(node_idx3 + k * eye[node_comp], shape, periodic)[0]
        kk = key(vx)
        j = np.clip(np.searchsorted(tks, kk), 0, len(tks) - 1)
        hit = tks[j] == kk
        slot_vert[hit, k] = torder[j[hit]]
        if (~hit).any():
            p = torch.as_tensor(origin + vx[~hit] * dx3, device=dev)
            slot_medium[~hit, k] = _classify(slot_fn, p).cpu().numpy()
    return TensorMesh(media=list(media), frac=frac[:, inT], normal=nhat[inT], node_comp=node_comp,
                      node_idx3=node_idx3, vert_nodes=inv.reshape(nV, 3, 2), shift=shift.reshape(nV, 3, 2, 3),
                      slot_vert=slot_vert, slot_medium=slot_medium, shape=shape, vert_idx3=T)


__all__ = ["MESH_MODES", "OnSubstrate", "ParticleSlots", "planar_background", "vertex_tensor", "voxelise"]
