"""The FDTD engine on MI355X: a drop-in for lightsmith's ``Engine`` (engine.py:230-719).

The constructor takes the reference's arguments (shape, dx, dt, single, inv_eps, groups, pml, sources,
monitors, dft_stride, periodic, bloch, offdiag, time_monitors, mirror, widths) and hands the time stepping
to the HIP library ``_lsf.so``: one H kernel and one E kernel per step (curl + CPML fused, the pure-ADE
update fused into the E kernel), plus sparse kernels for sources, interface groups and monitors.
Host set-up runs once in NumPy (float64) and uploads float32 tables.

Layout: each field component is stored z-major with x fastest (element (c, i, j, k) of the reference's
E[3, N0, N1, N2] at c*N + (k*N1 + j)*N0 + i). ``E``/``H`` are views in the reference's logical shape
(3, N0, N1, N2); ``E_flat``/``H_flat`` are the native flat arrays; ``native_index`` maps the reference's
flat indices (engine.py:103-105) to native ones. Periodic ghost planes are read through the Bloch phase
inside the kernels and are not materialised in the arrays.

Precision: float32 fields (complex64 under a Bloch phase), as the reference's GPU path (single=True).
Not supported here: mirror planes and graded grids (the reference's symmetry.py and grid.py paths).
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass, field

import numpy as np
import torch

from .. import _lib
from .._lib import (PH_ADVANCE_E, PH_ADVANCE_H, PH_CORRECT_E, PH_CORRECT_H, PH_GROUPS, PH_OFFDIAG, check, ptr)
from .media import Medium, pole_coefficients

__all__ = ["Engine", "PMLSpec", "DFTMonitor", "TimeMonitor", "DispersiveGroup", "OffDiagonal", "flat_index",
           "e_region", "h_region"]


def flat_index(c, i, j, k, shape):
    """The reference's flat index into E[3, N0, N1, N2] (engine.py:103-105)."""
    n0, n1, n2 = shape
    return ((np.asarray(c) * n0 + np.asarray(i)) * n1 + np.asarray(j)) * n2 + np.asarray(k)


def e_region(c: int, shape, periodic=(False, False, False)):
    """Stepped index box of E_c (engine.py:47-60, no mirror planes)."""
    return tuple(slice(1, n - 1) if (periodic[a] or a != c) else slice(0, n - 1) for a, n in enumerate(shape))


def h_region(c: int, shape, periodic=(False, False, False)):
    """Stepped index box of H_c (engine.py:63-66, no mirror planes)."""
    return tuple(slice(1, n - 1) if periodic[a] else (slice(0, n) if a == c else slice(0, n - 1))
                 for a, n in enumerate(shape))


@dataclass
class PMLSpec:
    """engine.py:184-191"""

    cells: tuple
    m: float = 3.0
    kappa_max: float = 1.0
    alpha_max: float = 0.0
    eps_ref: float = 1.0
    eps_ref_high: float | None = None


@dataclass
class DispersiveGroup:
    """engine.py:129-137: E nodes in D-form; pure when ``w`` is None (one medium)."""

    idx: np.ndarray
    media: list
    frac: np.ndarray
    w: np.ndarray | None


@dataclass
class OffDiagonal:
    """engine.py:117-126"""

    idx3: np.ndarray
    comp: np.ndarray
    coef: np.ndarray


@dataclass
class DFTMonitor:
    """engine.py:194-210: running DFT of E at ``e_idx`` and of (H[h_idx1] + H[h_idx2]) / 2."""

    omega: np.ndarray
    e_idx: np.ndarray
    h_idx1: np.ndarray | None = None
    h_idx2: np.ndarray | None = None
    energy: int = 0
    source: object | None = None
    acc_e: object = field(default=None, repr=False)
    acc_h: object = field(default=None, repr=False)

    def results(self):
        e = self.acc_e.cpu().numpy().astype(complex)
        h = None if self.acc_h is None else self.acc_h.cpu().numpy().astype(complex)
        return e, h


@dataclass
class TensorMesh:
    """engine.py:140-181: interface nodes updated with the smoothed 1/eps tensor applied at grid vertices."""

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


def _is_tensor(g) -> bool:
    """A TensorMesh of this package or of the reference (the same fields)."""
    return all(hasattr(g, a) for a in ("slot_vert", "vert_nodes", "node_comp", "normal"))


@dataclass
class TimeMonitor:
    """engine.py:213-224"""

    e_idx: np.ndarray
    n_max: int
    buf: object = field(default=None, repr=False)
    pos: object = field(default=None, repr=False)

    def results(self, n: int) -> np.ndarray:
        """(channels, samples) for the first n samples. The kernel drops samples past n_max and keeps counting,
        so an overflow raises here, as the reference's indexed store does at the sample (engine.py:590-592)."""
        taken = int(self.pos.item())
        if taken > int(self.n_max):
            raise IndexError(f"time monitor overflow: {taken} samples taken, buffer holds n_max={self.n_max}")
        return self.buf.cpu().numpy()[:n].T


def cpml_profile(spec: PMLSpec, n_lo: int, n_hi: int, N: int, h: float, dt: float, u: np.ndarray, high_eps=None):
    """(b, c, kappa) of the CPML grading at positions u (engine.py:358-371): polynomial order m,
    sigma_max = 0.8 (m + 1) / (sqrt(eps_ref) h), CFS alpha falling linearly into the layer."""
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


def filling_tiling(shapes, want: int = 16384) -> tuple:
    """A launch tiling (1 row per lane, 4 wavefronts per block, x cells per row, z chunk) for grids of these shapes
    (reference order: axis 0 along the rows, axis 2 marched) stepped together: 32-cell rows when no grid is wider,
    else 64, and the longest z chunk of 8, 4 or 2 planes that still gives the bulk kernels ``want`` row marches
    (16 per SIMD on 256 CUs). Small grids fill the GPU only with short z marches. Measured per step on the
    throughput scene at chunks 8 / 4 / 2 (us): 96^3 54.3 / 36.7 / 31.9, 128^3 61.3 / 49.8 / 50.1, 160^3
    87.9 / 87.3 / 88.3, 192^3 175.9 / 166.6 / 167.4, 256^3 378.9 / 373.1 / 384.6; 320^3 is fastest at 8."""
    shapes = [tuple(int(v) for v in s) for s in shapes]
    xw = 32 if max(s[0] for s in shapes) <= 32 else 64
    for zc in (8, 4, 2):
        if sum(-(-s[0] // xw) * s[1] * -(-s[2] // zc) for s in shapes) >= want:
            return (1, 4, xw, zc)
    return (1, 4, xw, 2)


def to_native(flat, shape) -> np.ndarray:
    """Native flat index (c*N + (k*N1 + j)*N0 + i) of reference flat indices ((c*N0 + i)*N1 + j)*N2 + k."""
    n0, n1, n2 = shape
    flat = np.asarray(flat, dtype=np.int64)
    k = flat % n2
    rest = flat // n2
    j = rest % n1
    rest //= n1
    i = rest % n0
    c = rest // n0
    return c * (n0 * n1 * n2) + (k * n1 + j) * n0 + i


def to_reference(native, shape) -> np.ndarray:
    """Inverse of to_native."""
    n0, n1, n2 = shape
    N = n0 * n1 * n2
    native = np.asarray(native, dtype=np.int64)
    c, cell = np.divmod(native, N)
    i = cell % n0
    rest = cell // n0
    j = rest % n1
    k = rest // n1
    return ((c * n0 + i) * n1 + j) * n2 + k


CHUNK = 32  # x cells per chunk of the D-form slot table (the narrowest wavefront row of the E kernel)


def ade_slot_key(nodes, shape) -> np.ndarray:
    """Slot order of D-form nodes: plane k, row j, 32-cell x chunk, component, then x -- a wavefront row's pole
    pairs of all three components form one contiguous block."""
    n0, n1, n2 = shape
    N = n0 * n1 * n2
    nchunk = (n0 + CHUNK - 1) // CHUNK
    c, cell = np.divmod(np.asarray(nodes, dtype=np.int64), N)
    x = cell % n0
    return (((cell // n0) * nchunk + x // CHUNK) * 3 + c) * CHUNK + x % CHUNK


def ade_chunk_base(nodes, shape) -> np.ndarray:
    """First compact slot of each (k, j, x // 32, component) for D-form nodes in slot order (ade_slot_key): the E
    kernel finds a node's slot as the base of its chunk and component + (D-form lanes below it in that chunk)."""
    n0, n1, n2 = shape
    nchunk = (n0 + CHUNK - 1) // CHUNK
    keys = ade_slot_key(nodes, shape) // CHUNK
    return np.searchsorted(keys, np.arange(n1 * n2 * nchunk * 3), side="left").astype(np.int32)


def encode_material(cb: np.ndarray):
    """(ids uint32 [N], table float32 [256]) with one 8-bit table index per component in bytes 0..2, or None
    when the (3, N) coefficients hold more than 256 distinct values."""
    flat = cb.reshape(-1)
    try:
        import pandas as pd  # hash-based unique: O(N)
        uniq = pd.unique(flat)
    except ImportError:  # pragma: no cover
        uniq = np.unique(flat)
    if len(uniq) > 256:
        return None, len(uniq)
    table = np.sort(np.asarray(uniq, dtype=np.float32))
    ids = np.searchsorted(table, cb).astype(np.uint32)
    packed = ids[0] | (ids[1] << 8) | (ids[2] << 16)
    tab = np.zeros(256, np.float32)
    tab[:len(table)] = table
    return (packed.astype(np.uint32), tab), len(uniq)


def pml_axis_profile(spec, shape, dt: float, spacing, d: int, field_: int):
    """(profile float64 [N_d, 4] = {b, c, 1/kappa, slab index or -1}, slab extent, has_kappa) of the CPML on
    axis d for E (field_ 0: stepped nodes 1..N-2 along d) or H (1: half-integer positions 0..N-2), the slab
    ranges of engine.py:388-395."""
    n_lo, n_hi = (int(v) for v in spec.cells[d])
    N, h = shape[d], spacing[d]
    start, stop, half = (1, N - 1, 0.0) if field_ == 0 else (0, N - 1, 0.5)
    ranges = []
    if n_lo:
        ranges.append((start, min(stop, n_lo + 1)))
    if n_hi:
        ranges.append((max(start, N - 2 - n_hi), stop))
    prof = np.zeros((N, 4))
    prof[:, 0], prof[:, 2], prof[:, 3] = 1.0, 1.0, -1.0
    ns, has_k = 0, False
    for a, b in ranges:
        if b <= a:
            continue
        u = np.arange(a, b) + half
        bb, cc, kk = cpml_profile(spec, n_lo, n_hi, N, h, dt, u)
        prof[a:b, 0], prof[a:b, 1], prof[a:b, 2] = bb, cc, 1.0 / kk
        prof[a:b, 3] = ns + np.arange(b - a)
        has_k |= bool(np.any(kk > 1))
        ns += b - a
    return prof, ns, has_k


def _device_of(device) -> torch.device:
    if hasattr(device, "native_device"):
        device = device.native_device()
    elif hasattr(device, "index") and hasattr(device, "framework"):
        device = getattr(device, "name", "cuda:0")
    dev = torch.device(device if device is not None else "cuda:0")
    if dev.type != "cuda":
        raise ValueError(f"lightsmith_amd runs on a GPU (torch 'cuda' = HIP); got device {dev}")
    return dev


class Engine:
    """FDTD engine whose stepping runs in the HIP library (engine.py:230-315 constructor contract)."""

    # launch tiling (rows per lane, wavefronts per block, x cells per wavefront row, z planes per block; the E
    # kernel of an engine with pole state marches twice the z planes); see lsf_set_tiling
    tiling = (1, 4, 64, 8)
    # launch the tensor nodes' D update with the E-side table correction when their nodes are disjoint
    fuse_sparse = True
    # default of the `fused` argument: step eligible engines (real fields, no periodic axes, one row per lane,
    # 64-cell rows, no H-phase operations) with the single-pass kernel over a second buffer set
    single_pass = False

    def __init__(self, device=None, *, shape, dx, dt: float, single: bool = True, inv_eps: np.ndarray,
                 groups=(), pml: PMLSpec | None = None, sources=(), monitors=(), dft_stride: int = 1,
                 periodic=(False, False, False), bloch=(0.0, 0.0, 0.0), offdiag: OffDiagonal | None = None,
                 time_monitors=(), mirror=(0, 0, 0), widths=None, material: str = "auto", fused: bool | None = None,
                 _sub: bool = False,
                 _counter=None):
        if not single:
            raise ValueError("the HIP engine steps float32 fields (single=True); double precision runs on the CPU")
        if any(mirror):
            raise ValueError("mirror planes are not supported by the HIP engine (run symmetry='off')")
        if widths is not None and any(w is not None for w in widths):
            raise ValueError("graded grids are not supported by the HIP engine (grid='uniform')")
        self.lib = _lib.load()
        self.device = device
        self._dev = _device_of(device)
        self.xp = torch
        self.shape = tuple(int(n) for n in shape)
        self.dt, self.single = float(dt), True
        self.spacing = tuple(float(v) for v in np.broadcast_to(np.asarray(dx, dtype=float), (3,)))
        self.uniform = len(set(self.spacing)) == 1
        self.dx = self.spacing[0]
        self.periodic = tuple(bool(p) for p in periodic)
        self.bloch = tuple(float(k) for k in bloch)
        if any(k and not p for k, p in zip(self.bloch, self.periodic)):
            raise ValueError("a Bloch wavevector needs a periodic axis")
        self.complex = any(self.bloch)
        self.mirror = (0, 0, 0)
        self.stride = int(dft_stride)
        self.rdt = torch.float32
        self.cdt = torch.complex64
        self.fdt = self.cdt if self.complex else self.rdt
        n0, n1, n2 = self.shape
        self.N = n0 * n1 * n2
        self._keep = []  # device tables whose lifetime is the engine's
        self._state = []  # device arrays the stepping writes (psi, pole and group state, source lines)
        self._subs = []  # sub-engines stepped inside this one
        self._held = 0  # batches stepping this engine
        # single-pass stepping (lsf_set_fused): a second buffer set for real non-periodic grids
        self._two_sets = (fused if fused is not None else Engine.single_pass) and not self.complex and not any(self.periodic) \
            and self.tiling[0] == 1 and self.tiling[2] == 64 and not _sub
        nsets = 2 if self._two_sets else 1
        self._E_bufs = [torch.zeros(3 * self.N, dtype=self.fdt, device=self._dev) for _ in range(nsets)]
        self._H_bufs = [torch.zeros(3 * self.N, dtype=self.fdt, device=self._dev) for _ in range(nsets)]
        view = lambda t: t.view(3, n2, n1, n0).permute(0, 3, 2, 1)  # noqa: E731
        self._E_views = [view(t) for t in self._E_bufs]
        self._H_views = [view(t) for t in self._H_bufs]
        self._psiH_blocks = []  # (block, size): the H-side psi arrays of an axis, both sets
        self._counter = _counter if _counter is not None else torch.zeros(2, dtype=torch.int64, device=self._dev)
        self.n = self._counter[0:1]
        self.n_host = 0
        self.RE = [e_region(c, self.shape, self.periodic) for c in range(3)]
        self.RH = [h_region(c, self.shape, self.periodic) for c in range(3)]
        self.cdx = self.dt / self.dx if self.uniform else self.dt  # engine.py:275
        self.cH = self.cdx

        g = _lib.LsfGrid()
        g.shape[:] = [n0, n1, n2]
        g.complex_fields = int(self.complex)
        g.periodic[:] = [int(p) for p in self.periodic]
        g.spacing[:] = list(self.spacing)
        g.dt = self.dt
        for a in range(3):
            kl = self.bloch[a] * (self.shape[a] - 2) * self.spacing[a] if self.periodic[a] else 0.0
            g.bloch_phase[a][0], g.bloch_phase[a][1] = np.cos(kl), np.sin(kl)
        g.dev_E, g.dev_H, g.dev_counter = ptr(self._E_bufs[0]), ptr(self._H_bufs[0]), ptr(self._counter)
        g.tick = 0 if _sub else 1
        handle = ctypes.c_void_p()
        check(self.lib.lsf_create(ctypes.byref(g), ctypes.byref(handle)), "lsf_create")
        self._h = handle
        self.tiling = tuple(int(v) for v in self.tiling)  # this engine's, whatever the class default becomes
        ry, waves, xw, zc = self.tiling
        check(self.lib.lsf_set_tiling(self._h, ry, waves, xw, zc), "lsf_set_tiling")
        self._e_table_nodes, self._tensor_nodes = [], []  # native targets, for the fused launch below

        inv_eps = np.asarray(inv_eps, dtype=float)
        if inv_eps.shape != (3,) + self.shape:
            raise ValueError(f"inv_eps must have shape {(3,) + self.shape}, got {inv_eps.shape}")
        groups = list(groups)
        tensors = [g for g in groups if _is_tensor(g)]
        dgroups = [g for g in groups if not _is_tensor(g)]
        for grp in dgroups:
            if not isinstance(grp, DispersiveGroup):
                raise ValueError(f"unsupported dispersive group type {type(grp).__name__}")
        self._build_ade(dgroups)
        self._set_material(inv_eps, material)
        self._build_pml(pml or PMLSpec(cells=((0, 0),) * 3), inv_eps)
        self._build_mixed(dgroups)
        for tm in tensors:
            self._build_tensor(tm)
        self._build_offdiag(offdiag, groups)
        self.energy = torch.zeros(2, dtype=torch.float32, device=self._dev)
        check(self.lib.lsf_set_energy(self._h, ptr(self.energy)))
        self.sources = list(sources)
        for s in self.sources:
            if not hasattr(s, "bind_native"):
                raise TypeError(f"source {type(s).__name__} has no HIP binding (bind_native)")
            s.bind(self)
        self.monitors = list(monitors)
        for m in self.monitors:
            self._bind_monitor(m)
        self.time_monitors = list(time_monitors)
        for m in self.time_monitors:
            self._bind_time_monitor(m)
        # the tensor nodes' D update rides on the E-side table launch when no table target is a tensor node
        self.fused_table_tensor = False
        if self.fuse_sparse and self._tensor_nodes and self._e_table_nodes:
            hit = np.intersect1d(np.concatenate(self._e_table_nodes), np.concatenate(self._tensor_nodes))
            self.fused_table_tensor = not hit.size and self.lib.lsf_fuse_table_tensor(self._h) == 0
        self._graph_stream = None
        self._graphed = False
        if self._two_sets:
            arr = (ctypes.c_void_p * 3)(*[None if t is None else t.data_ptr() for t in self._psiH_second])
            check(self.lib.lsf_set_fused(self._h, ptr(self._E_bufs[1]), ptr(self._H_bufs[1]), arr), "lsf_set_fused")

    def set_tiling(self, rows_per_lane: int, waves: int, x_cells: int, z_chunk: int):
        """Launch tiling of this engine's bulk kernels (lsf_set_tiling); before a graph is captured."""
        if self._graphed:
            raise ValueError("release the captured graph before changing the tiling")
        check(self.lib.lsf_set_tiling(self._h, int(rows_per_lane), int(waves), int(x_cells), int(z_chunk)),
              "lsf_set_tiling")
        self.tiling = (int(rows_per_lane), int(waves), int(x_cells), int(z_chunk))

    # --- fields: the current buffer set ------------------------------------------------------
    @property
    def parity(self) -> int:
        """Which buffer set holds the current fields (the single-pass step alternates them)."""
        return int(self.lib.lsf_parity(self._h)) if self._two_sets else 0

    @property
    def E_flat(self) -> torch.Tensor:
        return self._E_bufs[self.parity]

    @property
    def H_flat(self) -> torch.Tensor:
        return self._H_bufs[self.parity]

    @property
    def E(self) -> torch.Tensor:
        return self._E_views[self.parity]

    @property
    def H(self) -> torch.Tensor:
        return self._H_views[self.parity]

    # --- helpers ---------------------------------------------------------------------------
    def native_index(self, flat) -> np.ndarray:
        """Native flat index of reference flat indices into the stacked (3, N0, N1, N2) arrays."""
        return to_native(flat, self.shape)

    def reference_index(self, native) -> np.ndarray:
        """Inverse of native_index."""
        return to_reference(native, self.shape)

    def to_device(self, a, dtype=None):
        """Upload a NumPy array (kept alive for the engine's lifetime)."""
        a = np.asarray(a)
        if dtype is None:
            dtype = {np.dtype(np.int64): torch.int64, np.dtype(np.int32): torch.int32,
                     np.dtype(np.float64): torch.float32, np.dtype(np.float32): torch.float32,
                     np.dtype(np.complex128): torch.complex64, np.dtype(np.complex64): torch.complex64,
                     np.dtype(np.uint32): torch.int32}[a.dtype]
        if a.dtype == np.uint32:
            a = a.view(np.int32)
        t = torch.as_tensor(np.ascontiguousarray(a)).to(device=self._dev, dtype=dtype)
        self._keep.append(t)
        return t

    def zeros(self, n, complex_=None):
        """A zeroed device state array kept for the engine's lifetime."""
        t = torch.zeros(int(n), dtype=self.fdt if complex_ is None else (self.cdt if complex_ else self.rdt),
                        device=self._dev)
        self._keep.append(t)
        self._state.append(t)
        return t

    def field_values(self, vals) -> torch.Tensor:
        """Upload coefficients of the field type (complex under a Bloch phase)."""
        vals = np.asarray(vals)
        if self.complex:
            return self.to_device(vals.astype(np.complex128), torch.complex64)
        if np.iscomplexobj(vals):
            vals = vals.real
        return self.to_device(vals.astype(np.float64), torch.float32)

    @staticmethod
    def check_index(idx, n: int, what: str):
        """Every index the kernels dereference is validated here, on the host, before upload."""
        idx = np.asarray(idx)
        if idx.size and (idx.min() < 0 or idx.max() >= n):
            raise ValueError(f"{what}: index out of range [0, {n})")

    def _stream(self):
        return ctypes.c_void_p(torch.cuda.current_stream(self._dev).cuda_stream)

    # --- material and ADE --------------------------------------------------------------------
    def _build_ade(self, groups):
        """Pure groups: one D-form state slot per node, ordered by native index; chunk bases for the
        E kernel's ballot indexing; per-medium pole coefficients (engine.py:722-742)."""
        pure = [g for g in groups if g.w is None]
        self.ade_media: list[Medium] = []
        nodes, meds = [], []
        for g in pure:
            (m,) = g.media
            if not m.dispersive:
                raise ValueError("a pure D-form group needs a dispersive medium")
            if m not in self.ade_media:
                self.ade_media.append(m)
            mi = self.ade_media.index(m)
            nat = self.native_index(g.idx)
            c, cell = np.divmod(nat, self.N)
            i = cell % self.shape[0]
            j = (cell // self.shape[0]) % self.shape[1]
            k = cell // (self.shape[0] * self.shape[1])
            st = np.zeros(len(nat), bool)
            for cc in range(3):
                R = self.RE[cc]
                sel = c == cc
                st[sel] = ((i[sel] >= R[0].start) & (i[sel] < R[0].stop) & (j[sel] >= R[1].start) &
                           (j[sel] < R[1].stop) & (k[sel] >= R[2].start) & (k[sel] < R[2].stop))
            nodes.append(nat[st])
            meds.append(np.full(int(st.sum()), mi, dtype=np.int32))
        self.ade_nodes = np.concatenate(nodes) if nodes else np.zeros(0, np.int64)
        self.ade_medium = np.concatenate(meds) if meds else np.zeros(0, np.int32)
        order = np.argsort(self.ade_nodes, kind="stable")
        self.ade_nodes, self.ade_medium = self.ade_nodes[order], self.ade_medium[order]
        if len(self.ade_nodes) and np.any(np.diff(self.ade_nodes) == 0):
            raise ValueError("an E node belongs to two dispersive groups")
        # pole-state slots follow ade_slot_key; ade_nodes stays sorted by native index for lookups
        self.ade_slot_order = np.argsort(ade_slot_key(self.ade_nodes, self.shape), kind="stable")
        n = len(self.ade_nodes)
        self.n_ade = n
        if not n:
            return
        max_p = max(len(m.poles) for m in self.ade_media)
        if max_p > 4:
            raise ValueError("pure D-form media may have at most 4 poles on the HIP engine")
        self.ade_max_poles = max_p
        tab = np.zeros((len(self.ade_media), 2 + 3 * max_p))
        for q, m in enumerate(self.ade_media):
            tab[q, 0], tab[q, 1] = 1.0 / m.eps_inf, len(m.poles)
            for p, pole in enumerate(m.poles):
                tab[q, 2 + 3 * p:5 + 3 * p] = pole_coefficients(pole, self.dt)
        chunk_base = ade_chunk_base(self.ade_nodes[self.ade_slot_order], self.shape)
        self._ade_P = torch.zeros(max_p * n * 2, dtype=self.fdt, device=self._dev)
        self._keep.append(self._ade_P)
        self._state.append(self._ade_P)
        self._ade_tab = self.to_device(tab.astype(np.float32))
        self._ade_chunk = self.to_device(chunk_base)
        self._ade_inv_eps_inf = np.array([1.0 / m.eps_inf for m in self.ade_media])
        check(self.lib.lsf_set_ade(self._h, ptr(self._ade_chunk), ptr(self._ade_P), n, max_p, ptr(self._ade_tab),
                                   len(self.ade_media)), "lsf_set_ade")

    def correction_scale(self, native_targets):
        """Per-target factor of E corrections (device array, or None when every factor is 1): 1/eps_inf at pure
        D-form nodes, where the reference's correction enters through D before the group update."""
        if not self.n_ade:
            return None
        t = np.asarray(native_targets, dtype=np.int64)
        pos = np.searchsorted(self.ade_nodes, t)
        pos_c = np.minimum(pos, self.n_ade - 1)
        hit = (pos < self.n_ade) & (self.ade_nodes[pos_c] == t)
        if not hit.any():
            return None
        return self.to_device(np.where(hit, self._ade_inv_eps_inf[self.ade_medium[pos_c]], 1.0).astype(np.float32))

    def _set_material(self, inv_eps, mode: str):
        """E coefficients dt/dx * inv_eps per node (engine.py:299), D-form nodes encoded -(medium + 1).
        Stored as 8-bit indices into a 256-entry table when the grid has at most 256 distinct values."""
        n0, n1, n2 = self.shape
        cb = (self.cdx * inv_eps).astype(np.float32).transpose(0, 3, 2, 1).reshape(3, -1)  # native order
        flat = cb.reshape(-1)
        if not np.all(np.isfinite(flat)) or flat.min() < 0:
            raise ValueError("inv_eps must be finite and non-negative (dispersive cells go in groups)")
        if self.n_ade:
            flat[self.ade_nodes] = -(self.ade_medium.astype(np.float32) + 1.0)
        self.material_mode = "dense"
        if mode in ("auto", "indexed"):
            enc, n_uniq = encode_material(cb)
            if enc is not None:
                packed, tab = enc
                self._mat_ids = self.to_device(packed)
                self._mat_tab = self.to_device(tab)
                check(self.lib.lsf_set_material_indexed(self._h, ptr(self._mat_ids), ptr(self._mat_tab)))
                self.material_mode = "indexed"
                return
            if mode == "indexed":
                raise ValueError(f"{n_uniq} distinct material coefficients: more than an 8-bit table holds")
        self._mat = self.to_device(flat.astype(np.float32))
        check(self.lib.lsf_set_material_dense(self._h, ptr(self._mat)))

    # --- CPML (engine.py:344-418) --------------------------------------------------------------
    def _build_pml(self, spec: PMLSpec, inv_eps):
        self.pml_cells = tuple(tuple(int(v) for v in c) for c in spec.cells)
        self._psiH_second = [None, None, None]
        self.n_psi = 0  # psi values allocated (slab extent x full extent of the other two axes)
        self.n_psi_stepped = 0  # psi values the update touches (slab extent x the component's stepped box)
        for d in range(3):
            n_lo, n_hi = self.pml_cells[d]
            if not (n_lo or n_hi):
                continue
            if self.periodic[d]:
                raise ValueError("a periodic axis cannot have a PML")
            for fld in (0, 1):
                prof, ns, has_k = pml_axis_profile(spec, self.shape, self.dt, self.spacing, d, fld)
                if not ns:
                    continue
                others = [self.shape[a] for a in range(3) if a != d]
                size = ns * others[0] * others[1]
                psis = [None, None, None]
                regions = self.RE if fld == 0 else self.RH
                # both components of an axis in one allocation (the single-pass kernel addresses them from one
                # base); the H side of a two-set engine holds both sets
                if fld == 1 and self._two_sets:
                    block = torch.zeros(4 * size, dtype=self.fdt, device=self._dev)
                    self._keep.append(block)
                    self._psiH_blocks.append((block, size))
                    self._state.append(len(self._psiH_blocks) - 1)  # resolved to the current set by parity
                else:
                    block = self.zeros(2 * size)
                comps = [c for c in range(3) if c != d]
                for c in range(3):
                    if c != d:
                        slot = comps.index(c)
                        psis[c] = block[slot * size:(slot + 1) * size]
                        self.n_psi += size
                        o1, o2 = (a for a in range(3) if a != d)
                        R = regions[c]
                        self.n_psi_stepped += ns * (R[o1].stop - R[o1].start) * (R[o2].stop - R[o2].start)
                tprof = self.to_device(prof.astype(np.float32))
                check(self.lib.lsf_set_pml(self._h, fld, d, ptr(tprof), ns, ptr(psis[0]), ptr(psis[1]),
                                           ptr(psis[2]), int(has_k)), "lsf_set_pml")
                if fld == 1 and self._two_sets:
                    self._psiH_second[d] = block[2 * size:]

    # --- mixed (interface) groups (engine.py:743-791) --------------------------------------------
    def _build_mixed(self, groups):
        for g in groups:
            if g.w is None:
                continue
            media = list(g.media)
            f = np.asarray(g.frac, dtype=float)
            S = len(media)
            maxp = max([len(m.poles) for m in media] + [1])
            nat = self.native_index(g.idx)
            self.check_index(nat, 3 * self.N, "mixed group index")
            h0 = sum(f[s] / m.eps_inf for s, m in enumerate(media) if not m.dispersive) + 0 * f[0]
            inv_a = 1.0 / sum(f[s] * m.eps_inf for s, m in enumerate(media))
            w = np.asarray(g.w, dtype=float)
            sm = np.zeros((S, 2 + 3 * maxp))
            for s, m in enumerate(media):
                sm[s, 0], sm[s, 1] = 1.0 / m.eps_inf, len(m.poles)
                for p, pole in enumerate(m.poles):
                    sm[s, 2 + 3 * p:5 + 3 * p] = pole_coefficients(pole, self.dt)
            n = len(nat)
            state = self.zeros((3 + S + 4 * S * maxp) * n)
            arrs = [self.to_device(nat), self.to_device(f.astype(np.float32)), self.to_device(w.astype(np.float32)),
                    self.to_device(np.asarray(h0, np.float32)), self.to_device(np.asarray(inv_a, np.float32)),
                    self.to_device(sm.astype(np.float32))]
            check(self.lib.lsf_add_mixed_group(self._h, n, S, maxp, *(ptr(a) for a in arrs), ptr(state),
                                               int(np.any(w > 0)), int(np.any(w < 1))), "lsf_add_mixed_group")

    # --- conformal vertex tensor (engine.py:839-923) ---------------------------------------------
    def _build_tensor(self, g):
        if getattr(g, "sign", None) is not None:
            raise ValueError("vertex tensors next to mirror planes are not supported (run symmetry='off')")
        nN, nV = len(g.node_comp), len(g.normal)
        nat = self.native_index(g.idx)
        self.check_index(nat, 3 * self.N, "tensor node index")
        media = list(g.media)
        S = len(media)
        maxp = max([len(m.poles) for m in media] + [1])
        sv = np.asarray(g.slot_vert, dtype=np.int64)
        slot_v = np.concatenate([sv[:, 0], sv[:, 1]])  # slot k * nN + node
        slot_m = np.concatenate([g.slot_medium[:, 0], g.slot_medium[:, 1]])
        frac = np.asarray(g.frac, dtype=float)
        F = np.zeros((S, 2 * nN))
        tensor = slot_v >= 0
        F[:, tensor] = frac[:, slot_v[tensor]]
        F[slot_m[~tensor], np.flatnonzero(~tensor)] = 1.0
        h0 = sum(F[s] / m.eps_inf for s, m in enumerate(media) if not m.dispersive) + 0 * F[0]
        inv_a = 1.0 / sum(F[s] * m.eps_inf for s, m in enumerate(media))
        sm = np.zeros((S, 2 + 3 * maxp))
        for s, m in enumerate(media):
            sm[s, 0], sm[s, 1] = 1.0 / m.eps_inf, len(m.poles)
            for p, pole in enumerate(m.poles):
                sm[s, 2 + 3 * p:5 + 3 * p] = pole_coefficients(pole, self.dt)
        vph = None
        if self.complex:
            L = np.array([(n - 2) * d for n, d in zip(self.shape, self.spacing)])
            ph = np.exp(1j * (np.asarray(g.shift) * (np.asarray(self.bloch) * L)).sum(-1))  # (nV, 3, 2)
            vph = self.to_device(ph.reshape(-1).astype(np.complex128))
        state = self.zeros((6 + 2 * S + 8 * S * maxp) * nN)
        arrs = dict(nodes=self.to_device(nat), comp=self.to_device(np.asarray(g.node_comp, np.int32)),
                    vn=self.to_device(np.asarray(g.vert_nodes, np.int32).reshape(-1)),
                    normal=self.to_device(np.asarray(g.normal, np.float32).reshape(-1)),
                    slot_v=self.to_device(slot_v.astype(np.int32)), f=self.to_device(F.astype(np.float32)),
                    h0=self.to_device(np.asarray(h0, np.float32)), inv_a=self.to_device(inv_a.astype(np.float32)),
                    sm=self.to_device(sm.astype(np.float32)))
        check(self.lib.lsf_add_tensor_group(self._h, nN, ptr(arrs["nodes"]), ptr(arrs["comp"]), nV, ptr(arrs["vn"]),
                                            ptr(vph), ptr(arrs["normal"]), ptr(arrs["slot_v"]), S, maxp,
                                            ptr(arrs["f"]), ptr(arrs["h0"]), ptr(arrs["inv_a"]), ptr(arrs["sm"]),
                                            ptr(state)), "lsf_add_tensor_group")
        self.n_tensor_nodes = getattr(self, "n_tensor_nodes", 0) + nN
        self._tensor_nodes.append(nat)

    # --- off-diagonal tensor terms (engine.py:421-518, uniform grid) ------------------------------
    def _build_offdiag(self, od, groups):
        self.od = None
        if od is None or not len(od.comp):
            return
        shape = self.shape
        nflat = 3 * self.N
        idx3 = np.asarray(od.idx3, dtype=np.int64)
        comp = np.asarray(od.comp, dtype=np.int64)
        coef = np.asarray(od.coef, dtype=float)
        eye = np.eye(3, dtype=np.int64)
        tc, ti, sc, si, fwd, bwd = [], [], [], [], [], []
        for d in range(3):
            for corner in range(4):
                q = idx3 + eye[comp] * (corner & 1)
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
                fwd.append(coef[ok, d] / 8 * phase[ok])
                bwd.append(np.conj(coef[ok, d] * phase[ok]) / 8)
        tc, ti, sc, si = (np.concatenate(x) for x in (tc, ti, sc, si))
        v = np.concatenate([np.concatenate(fwd), np.concatenate(bwd)])
        tc, sc = np.concatenate([tc, sc]), np.concatenate([sc, tc])
        ti, si = np.concatenate([ti, si]), np.concatenate([si, ti])
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
        hh = np.zeros((len(targets), 4 * width), dtype=np.int64)
        ww = np.zeros((len(targets), 4 * width), dtype=complex)
        dd, rest = np.divmod(s_, self.N)
        q = np.stack(np.unravel_index(rest, shape), axis=1)
        d1, d2 = (dd + 1) % 3, (dd + 2) % 3
        inv_d = 1.0 / np.asarray(self.spacing)
        # curl_d H at q = (H_d2[q] - H_d2[q - e_d1]) / dx_d1 - (H_d1[q] - H_d1[q - e_d2]) / dx_d2
        for k, (cmp, back, sgn, ax) in enumerate(((d2, None, 1.0, d1), (d2, d1, -1.0, d1),
                                                  (d1, None, -1.0, d2), (d1, d2, 1.0, d2))):
            p = q if back is None else q - eye[back]
            hh[row, 4 * slot + k] = flat_index(cmp, p[:, 0], p[:, 1], p[:, 2], shape)
            ww[row, 4 * slot + k] = self.dt * sgn * inv_d[ax] * v
        self.od = (targets, hh, ww)
        tgt = self.native_index(targets)
        src = self.native_index(hh.reshape(-1))
        self._od = [self.to_device(tgt), self.to_device(src), self.field_values(ww.reshape(-1))]
        check(self.lib.lsf_add_gather(self._h, PH_OFFDIAG, 0, len(tgt), 4 * width, ptr(self._od[0]),
                                      None, ptr(self._od[1]), ptr(self._od[2]), None),
              "lsf_add_gather (off-diagonal)")

    # --- sparse work used by sources ---------------------------------------------------------------
    def add_gather(self, phase: int, field_: str, targets, src_tensor, src_idx, weights):
        """Rows F[target] += sum_k w[k] * src[src_idx[k]]; duplicate targets are merged into one row."""
        targets = np.asarray(targets, dtype=np.int64)
        src_idx = np.asarray(src_idx, dtype=np.int64).reshape(len(targets), -1)
        weights = np.asarray(weights).reshape(len(targets), -1)
        uniq, inv, counts = np.unique(targets, return_inverse=True, return_counts=True)
        taps = src_idx.shape[1]
        width = int(counts.max()) * taps
        order = np.argsort(inv, kind="stable")
        slot = np.arange(len(targets)) - np.repeat(np.cumsum(counts) - counts, counts)
        S = np.zeros((len(uniq), width), dtype=np.int64)
        W = np.zeros((len(uniq), width), dtype=weights.dtype)
        rows = inv[order]
        for t in range(taps):
            S[rows, slot * taps + t] = src_idx[order, t]
            W[rows, slot * taps + t] = weights[order, t]
        fld = 0 if field_ == "E" else 1
        self.check_index(uniq, 3 * self.N, "gather target")
        self.check_index(S, src_tensor.numel(), "gather source")
        sc = self.correction_scale(uniq) if fld == 0 else None
        arrs = [self.to_device(uniq), self.to_device(S.reshape(-1)), self.field_values(W.reshape(-1))]
        check(self.lib.lsf_add_gather(self._h, phase, fld, len(uniq), width, ptr(arrs[0]), ptr(src_tensor),
                                      ptr(arrs[1]), ptr(arrs[2]), ptr(sc)), "lsf_add_gather")

    def add_modal(self, phase: int, field_: str, targets, m, table):
        """F[target[r]] += sum_q m[r, q] table[n, q] (targets must be distinct)."""
        targets = np.asarray(targets, dtype=np.int64)
        if len(np.unique(targets)) != len(targets):
            raise ValueError("modal source targets must be distinct")
        m = np.asarray(m).reshape(len(targets), -1)
        table = np.asarray(table, dtype=np.float64).reshape(-1, m.shape[1])
        self.check_index(targets, 3 * self.N, "modal target")
        fld = 0 if field_ == "E" else 1
        sc = self.correction_scale(targets) if fld == 0 else None
        arrs = [self.to_device(targets), self.field_values(m.reshape(-1)), self.to_device(table.astype(np.float32))]
        check(self.lib.lsf_add_modal(self._h, phase, fld, len(targets), m.shape[1], ptr(arrs[0]), ptr(arrs[1]),
                                     ptr(arrs[2]), table.shape[0], ptr(sc)), "lsf_add_modal")

    def add_modal_to(self, phase: int, field_: str, targets, m, table):
        self.add_modal(phase, field_, targets, m, table)

    def add_table(self, phase: int, field_: str, nodes, T, m: int, cmax: int, base, i0, u0, coef):
        """Table-fed corrections, points grouped by target node (CSR)."""
        nodes = np.asarray(nodes, dtype=np.int64)
        order = np.argsort(nodes, kind="stable")
        uniq, counts = np.unique(nodes[order], return_counts=True)
        row_ptr = np.concatenate([[0], np.cumsum(counts)]).astype(np.int32)
        self.check_index(uniq, 3 * self.N, "table target")
        base, i0 = np.asarray(base, np.int64), np.asarray(i0, np.int64)
        # reads T[base + c + j], j = 0..5, with floor(i0 / m) <= c <= cmax at every step t >= 0
        if base.size and (i0.min() < 0 or (base + i0 // int(m)).min() < 0 or base.max() + int(cmax) + 6 > T.numel()):
            raise ValueError("table source: reads outside the tables")
        fld = 0 if field_ == "E" else 1
        if fld == 0 and phase == PH_CORRECT_E:
            self._e_table_nodes.append(uniq)
        sc = self.correction_scale(uniq) if fld == 0 else None
        arrs = [self.to_device(uniq), self.to_device(row_ptr), T,
                self.to_device(np.asarray(base, np.int64)[order]), self.to_device(np.asarray(i0, np.int64)[order]),
                self.to_device(np.asarray(u0, np.float32)[order]), self.to_device(np.asarray(coef, np.float32)[order])]
        check(self.lib.lsf_add_table(self._h, phase, fld, len(uniq), ptr(arrs[0]), ptr(arrs[1]), ptr(arrs[2]), int(m),
                                     int(cmax), ptr(arrs[3]), ptr(arrs[4]), ptr(arrs[5]), ptr(arrs[6]), ptr(sc)),
              "lsf_add_table")

    # --- monitors (engine.py:556-592) ----------------------------------------------------------
    def _bind_monitor(self, m: DFTMonitor):
        omega = np.asarray(m.omega, dtype=float)
        src = m.source
        if src is None:
            e_idx = self.native_index(m.e_idx)
            h1 = None if m.h_idx1 is None else self.native_index(m.h_idx1)
            h2 = None if m.h_idx2 is None else self.native_index(m.h_idx2)
            esrc = hsrc = None
            cx = int(self.complex)
            n_e = n_h = 3 * self.N
        else:
            esrc, hsrc, cx, remap = src.monitor_arrays()
            e_idx = remap(m.e_idx)
            h1 = None if m.h_idx1 is None else remap(m.h_idx1)
            h2 = None if m.h_idx2 is None else remap(m.h_idx2)
            n_e, n_h = esrc.numel(), hsrc.numel()
        self.check_index(e_idx, n_e, "monitor E index")
        if h1 is not None:
            self.check_index(h1, n_h, "monitor H index")
            self.check_index(h2, n_h, "monitor H index")
        m._omega = self.to_device(omega, torch.float64)
        m._e = self.to_device(e_idx)
        m.acc_e = torch.zeros((len(omega), len(e_idx)), dtype=torch.complex64, device=self._dev)
        m.acc_h = None
        if h1 is not None:
            m._h1, m._h2 = self.to_device(h1), self.to_device(h2)
            m.acc_h = torch.zeros((len(omega), len(h1)), dtype=torch.complex64, device=self._dev)
        else:
            m._h1 = m._h2 = None
        check(self.lib.lsf_add_dft_monitor(self._h, len(omega), len(e_idx), ptr(m._omega), ptr(m._e), ptr(m._h1),
                                           ptr(m._h2), ptr(m.acc_e), ptr(m.acc_h), ptr(esrc), ptr(hsrc), int(cx),
                                           int(m.energy) if m.acc_h is not None else 0), "lsf_add_dft_monitor")

    def _bind_time_monitor(self, m: TimeMonitor):
        e_idx = self.native_index(m.e_idx)
        self.check_index(e_idx, 3 * self.N, "time monitor index")
        m._e = self.to_device(e_idx)
        m.buf = torch.zeros((int(m.n_max), len(e_idx)), dtype=self.fdt, device=self._dev)
        m.pos = torch.zeros(1, dtype=torch.int64, device=self._dev)
        check(self.lib.lsf_add_time_monitor(self._h, len(e_idx), ptr(m._e), ptr(m.buf), int(m.n_max), ptr(m.pos)))

    # --- stepping (engine.py:665-719) ------------------------------------------------------------
    def step(self):
        check(self.lib.lsf_step(self._h, 1, self._stream()), "lsf_step")
        self.n_host += 1

    def steps(self, n: int):
        check(self.lib.lsf_step(self._h, int(n), self._stream()), "lsf_step")
        self.n_host += int(n)

    def sample(self):
        check(self.lib.lsf_sample(self._h, self._stream()), "lsf_sample")

    def chunk(self):
        check(self.lib.lsf_chunk(self._h, 1, self.stride, self._stream()), "lsf_chunk")
        self.n_host += self.stride

    def make_runner(self, use_graphs: bool):
        """(run_one_chunk, used_graphs): with graphs, one chunk is captured as a HIP graph on a side stream
        and replayed (engine.py:685-709; the warm-up chunk runs eagerly)."""
        if not use_graphs:
            return self.chunk, False
        stream = torch.cuda.Stream(device=self._dev)
        stream.wait_stream(torch.cuda.current_stream(self._dev))
        state = {"captured": False}
        self._graph_stream = stream

        def run():
            if not state["captured"]:
                self.chunk()  # the warm-up chunk: real steps, eager
                torch.cuda.current_stream(self._dev).synchronize()
                check(self.lib.lsf_graph_capture(self._h, self.stride, ctypes.c_void_p(stream.cuda_stream)),
                      "lsf_graph_capture")
                state["captured"] = True
                self._graphed = True
                return
            stream.wait_stream(torch.cuda.current_stream(self._dev))
            check(self.lib.lsf_graph_replay(self._h, 1, ctypes.c_void_p(stream.cuda_stream)), "lsf_graph_replay")
            torch.cuda.current_stream(self._dev).wait_stream(stream)
            self.n_host += self.stride

        return run, True

    def release(self):
        """Drop a captured graph once the run is over (engine.py:711-719)."""
        if self._graphed:
            check(self.lib.lsf_graph_reset(self._h), "lsf_graph_reset")
            self._graphed = False

    def add_subengine(self, sub: "Engine"):
        """Step `sub` (created with _sub=True on this engine's counter) inside this engine's H and E phases."""
        check(self.lib.lsf_add_subengine(self._h, sub._h), "lsf_add_subengine")
        self._subs.append(sub)

    def state_tensors(self) -> list:
        """Every device tensor the engine reads or writes while stepping (fields, counter, psi, pole and group
        state, source lines, monitor accumulators, sub-engines'), each once: a snapshot of these restores the
        run exactly."""
        out, seen = [], set()
        par = self.parity
        ts = [self.E_flat, self.H_flat, self._counter, self.energy]
        for t in self._state:  # an int names an H-side psi block of a two-set engine: its current half
            if isinstance(t, int):
                block, size = self._psiH_blocks[t]
                t = block[par * 2 * size:(par + 1) * 2 * size]
            ts.append(t)
        for m in self.monitors:
            ts += [getattr(m, "acc_e", None), getattr(m, "acc_h", None)]
        for m in self.time_monitors:
            ts += [getattr(m, "buf", None), getattr(m, "pos", None)]
        for sub in self._subs:
            ts += sub.state_tensors()
        for t in ts:
            if isinstance(t, torch.Tensor) and t.data_ptr() not in seen:
                seen.add(t.data_ptr())
                out.append(t)
        return out

    def close(self):
        if self._held:
            raise RuntimeError("the engine is stepped by a Batch: close the batch first")
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            torch.cuda.synchronize(self._dev)
            self.lib.lsf_destroy(h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            if not self._held:
                self.close()
        except Exception:  # noqa: BLE001 (interpreter shutdown)
            pass

    # --- diagnostics -------------------------------------------------------------------------------
    def bytes_per_step(self) -> dict:
        """Algorithmic HBM bytes of one step: fields read + written, 3 float32 coefficients per cell,
        psi read + written, D-form state (D, E_prev, P^n, P^{n-1} per pole) read + written, as counted for
        the reference's update; the kernels keep neither D nor E_prev, so they move less than this."""
        s = 8 if self.complex else 4
        n_ade = self.n_ade
        p = getattr(self, "ade_max_poles", 0)
        fields = self.N * (12 * s + 12)
        psi = self.n_psi_stepped * 2 * s
        ade = n_ade * (2 + 2 * p) * 2 * s
        return {"fields": fields, "psi": psi, "ade": ade, "total": fields + psi + ade}
