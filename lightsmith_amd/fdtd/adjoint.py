"""Gradients of DFT-monitor objectives with respect to a design region's permittivity, by the adjoint method.

The reference FDTD has no adjoint (docs/solvers/fdtd.md:314; its inverse design runs on FDFD,
inverse/__init__.py:14-40); this module adds one behind ``torch.autograd``:

    prob = AdjointProblem(device, shape=..., design=DesignRegion(...), monitor=DFTMonitor(...), ...)
    acc_e, acc_h = monitor_spectra(rho, prob)          # forward FDTD run, differentiable in rho
    loss = objective(acc_e, acc_h); loss.backward()    # rho.grad by one adjoint run

Density rho in [0, 1] per design voxel sets eps = eps_lo + rho (eps_hi - eps_lo) at the E_x, E_y and E_z nodes of
that cell index (staircase). The scene may hold pure D-form (Drude-Lorentz ADE) groups outside the design region;
the adjoint engine runs their transposed pole recursion (lsf_add_adjoint_ade). The forward engine records, per step, X = curl- H at the design nodes (a tape); the
adjoint engine (lsf_set_adjoint) runs the transposed update backwards from the monitor's adjoint sources and
accumulates dL/dcb = sum_n Ebar^{n+1} X^{n+1/2} per node; dL/drho follows from cb = (dt/dx) / eps(rho).

Memory: the tape holds 4 B per design node and step. When all steps fit in `tape_budget` bytes the adjoint needs
no recomputation. Otherwise the forward run keeps a checkpoint of its state every `checkpoint_every` steps, and
the adjoint pass recomputes each interval from its checkpoint (recording that interval's tape) just before
reversing it: memory stays at the checkpoints plus one interval of tape.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import torch

from .._lib import check, ptr
from .engine import DFTMonitor, Engine, PMLSpec, filling_tiling


@dataclass
class DesignRegion:
    """Cells [lo, hi) (index boxes per axis); eps = eps_lo + rho (eps_hi - eps_lo) at their E nodes."""

    lo: tuple
    hi: tuple
    eps_lo: float
    eps_hi: float

    @property
    def shape(self) -> tuple:
        return tuple(int(h - l) for l, h in zip(self.lo, self.hi))

    @property
    def n_voxels(self) -> int:
        return int(np.prod(self.shape))

    def nodes(self, shape) -> tuple[np.ndarray, np.ndarray]:
        """(native flat E indices, voxel index) of the 3 * n_voxels design nodes, component-major; voxels in
        C order over the box (i slowest, k fastest, as rho.reshape(shape))."""
        n0, n1, n2 = shape
        N = n0 * n1 * n2
        ii, jj, kk = np.meshgrid(*[np.arange(l, h) for l, h in zip(self.lo, self.hi)], indexing="ij")
        cell = ((kk * n1 + jj) * n0 + ii).reshape(-1)
        nodes = np.concatenate([c * N + cell for c in range(3)]).astype(np.int64)
        vox = np.tile(np.arange(self.n_voxels), 3)
        return nodes, vox


class _Rows:
    """Monitor points grouped by the grid node they touch, as CSR rows in a fixed order (node, then point): the
    adjoint source of points that share a node (edges and corners of a flux box) is summed per row by one thread,
    so gradients do not depend on scheduling."""

    def __init__(self, targets, points, dev):
        targets, points = np.asarray(targets, dtype=np.int64), np.asarray(points, dtype=np.int64)
        order = np.lexsort((points, targets))
        t, p = targets[order], points[order]
        uniq, start, count = np.unique(t, return_index=True, return_counts=True)
        self.n_rows = len(uniq)
        self.target = torch.as_tensor(uniq, device=dev)
        self.entry_target = torch.as_tensor(t, device=dev)
        self.point = torch.as_tensor(p, device=dev)
        self.row_ptr = torch.as_tensor(np.append(start, len(t)).astype(np.int32), device=dev)
        # padded (row, slot) -> entry map for a fixed-order row sum in torch (slot past the row: a zero entry)
        width = int(count.max())
        slot = np.arange(width)[None, :]
        pad = np.where(slot < count[:, None], start[:, None] + slot, len(t))
        self._pad = torch.as_tensor(pad, device=dev)

    def row_sum(self, per_point: torch.Tensor, per_entry_scale: torch.Tensor) -> torch.Tensor:
        """sum over each row's entries of per_entry_scale * per_point[point], in the rows' fixed order."""
        vals = per_point[self.point] * per_entry_scale
        vals = torch.cat([vals, vals.new_zeros(1)])
        return vals[self._pad].sum(1)


class AdjointProblem:
    """A real-field FDTD run (CPML, sources, one DFT monitor) whose design-region permittivity follows rho."""

    def __init__(self, device, *, shape, dx, dt: float, inv_eps: np.ndarray, pml: PMLSpec, sources, monitor: DFTMonitor,
                 design: DesignRegion, n_steps: int, dft_stride: int, checkpoint_every: int = 100,
                 tape_budget: int = 4 << 30, use_graphs: bool = True, groups=()):
        self.design = design
        self.shape = tuple(int(v) for v in shape)
        self.stride = int(dft_stride)
        self.n_steps = int(np.ceil(n_steps / self.stride)) * self.stride
        K = max(self.stride, int(np.ceil(checkpoint_every / self.stride)) * self.stride)
        self.K = K
        self.use_graphs = use_graphs
        for a in range(3):
            cells = pml.cells[a]
            if design.lo[a] < max(1, cells[0]) or design.hi[a] > self.shape[a] - max(1, cells[1]) - 1:
                raise ValueError("the design region must lie inside the CPML-free interior")
        groups = list(groups)
        if any(getattr(g, "w", None) is not None or hasattr(g, "slot_vert") for g in groups):
            raise ValueError("the adjoint transposes pure D-form groups only (no mixed or vertex-tensor groups)")
        self.fwd = Engine(device, shape=shape, dx=dx, dt=dt, single=True, inv_eps=inv_eps, groups=groups, pml=pml,
                          sources=sources, monitors=[monitor], dft_stride=self.stride, material="dense", fused=False)
        ry, waves, _, zc = filling_tiling([self.fwd.shape])
        tiling = (ry, waves, 64, zc)  # the adjoint kernels take 64-cell rows
        self.fwd.set_tiling(*tiling)
        self.monitor = monitor
        dev = self.fwd._dev
        self._dev = dev
        nodes, vox = design.nodes(self.shape)
        if self.fwd.n_ade:
            if np.isin(nodes, self.fwd.ade_nodes).any():
                raise ValueError("the design region must not hold D-form (dispersive) nodes")
            # the transposed CPML recursions read the adjoint field of their node; D-form nodes keep out of the slabs
            n0, n1, _ = self.shape
            cell = self.fwd.ade_nodes % self.fwd.N
            idx = (cell % n0, (cell // n0) % n1, cell // (n0 * n1))
            for a in range(3):
                lo_c, hi_c = pml.cells[a]
                if idx[a].min() < max(1, lo_c) or idx[a].max() > self.shape[a] - max(1, hi_c) - 1:
                    raise ValueError("D-form (dispersive) nodes must lie inside the CPML-free interior")
        self.n_nodes = len(nodes)
        self._nodes = torch.as_tensor(nodes, device=dev)
        self._vox = torch.as_tensor(vox, device=dev)
        self.cb_design = self.fwd._mat[self._nodes].clone()
        per_step = 4 * self.n_nodes
        self.store_all = per_step * self.n_steps <= tape_budget
        self.n_slots = self.n_steps if self.store_all else K
        self.tape = torch.zeros(self.n_slots * self.n_nodes, dtype=torch.float32, device=dev)
        self.base = torch.zeros(1, dtype=torch.int64, device=dev)
        lib = self.fwd.lib
        check(lib.lsf_add_tape(self.fwd._h, self.n_nodes, ptr(self._nodes), ptr(self.tape), self.n_slots,
                               ptr(self.base)), "lsf_add_tape")

        # the adjoint engine: same grid, material and CPML profiles; its own fields and chi arrays. Its material is
        # the forward one with the D-form nodes' code replaced by their coefficient dt/dx (inv_eps = 1 there);
        # set_density updates both at the design nodes
        self.adj = Engine(device, shape=shape, dx=dx, dt=dt, single=True, inv_eps=inv_eps, pml=pml, sources=[],
                          monitors=[], dft_stride=1, material="dense", fused=False)
        self.adj.set_tiling(*tiling)
        self._adj_mat = self.fwd._mat.clone()
        if self.fwd.n_ade:
            self._adj_mat[torch.as_tensor(self.fwd.ade_nodes, device=dev)] = float(self.fwd.cdx)
        check(lib.lsf_set_material_dense(self.adj._h, ptr(self._adj_mat)), "lsf_set_material_dense")
        check(lib.lsf_set_adjoint(self.adj._h, 1), "lsf_set_adjoint")
        if self.fwd.n_ade:
            f = self.fwd
            self._ade_nodes = torch.as_tensor(f.ade_nodes, device=dev)
            self._ade_medium = torch.as_tensor(f.ade_medium.astype(np.int32), device=dev)
            self._ade_state = self.adj.zeros(f.n_ade * 2 * f.ade_max_poles, complex_=False)
            self._ade_save = torch.zeros(f.n_ade, dtype=torch.float32, device=dev)
            check(lib.lsf_add_adjoint_ade(self.adj._h, f.n_ade, ptr(self._ade_nodes), ptr(self._ade_medium),
                                          len(f.ade_media), f.ade_max_poles, ptr(f._ade_tab), ptr(self._ade_state),
                                          ptr(self._ade_save)), "lsf_add_adjoint_ade")
        self.g = torch.zeros(self.n_nodes, dtype=torch.float64, device=dev)
        check(lib.lsf_add_adjoint_gradient(self.adj._h, self.n_nodes, ptr(self._nodes), ptr(self.cb_design),
                                           ptr(self.tape), self.n_slots, ptr(self.base), self.n_steps, ptr(self.g)),
              "lsf_add_adjoint_gradient")
        m = monitor
        nf = len(np.asarray(m.omega))
        e_nat = self.fwd.native_index(m.e_idx)
        self._e_rows = _Rows(e_nat, np.arange(len(e_nat)), dev)
        self._e_nodes = self._e_rows.target
        self._cb_e = self._adj_mat[self._e_rows.entry_target].clone()
        self.W_E = torch.zeros((nf, len(e_nat)), dtype=torch.complex64, device=dev)
        r = self._e_rows
        check(lib.lsf_add_adjoint_source(self.adj._h, 0, nf, len(e_nat), ptr(m._omega), float(dt), ptr(self.W_E),
                                         r.n_rows, ptr(r.row_ptr), ptr(r.target), ptr(r.point), ptr(self._cb_e),
                                         self.n_steps, self.stride), "lsf_add_adjoint_source")
        self.W_H = None
        if m.acc_h is not None:
            h1, h2 = self.fwd.native_index(m.h_idx1), self.fwd.native_index(m.h_idx2)
            n_h = len(h1)
            self._h_rows = _Rows(np.concatenate([h1, h2]), np.tile(np.arange(n_h), 2), dev)
            self.W_H = torch.zeros((nf, n_h), dtype=torch.complex64, device=dev)
            self._h_scale = torch.full((len(self._h_rows.entry_target),), -0.5 * self.fwd.cH, dtype=torch.float32,
                                       device=dev)
            r = self._h_rows
            check(lib.lsf_add_adjoint_source(self.adj._h, 1, nf, n_h, ptr(m._omega), float(dt), ptr(self.W_H),
                                             r.n_rows, ptr(r.row_ptr), ptr(r.target), ptr(r.point),
                                             ptr(self._h_scale), self.n_steps, self.stride), "lsf_add_adjoint_source")
        self._omega = torch.as_tensor(np.asarray(m.omega, dtype=float), device=dev)
        self.dt = float(dt)
        self._run = None
        self.checkpoints = []
        self.timings = {}
        self.runs = 0  # forward runs so far: the tape, checkpoints and design cb belong to the latest one

    def use_checkpoints(self, on: bool):
        """Switch between the stored tape (no recompute) and checkpoint recompute (one interval of tape in use);
        the tape allocated at construction must hold at least one checkpoint interval."""
        if not on and self.n_slots < self.n_steps:
            raise ValueError("the tape was sized for checkpoint recompute (tape_budget); it cannot hold every step")
        if on and self.n_slots < self.K:
            raise ValueError("the tape holds less than one checkpoint interval")
        self.store_all = not on

    # --- density -> material --------------------------------------------------------------------------
    def eps(self, rho: torch.Tensor) -> torch.Tensor:
        d = self.design
        return d.eps_lo + rho.to(torch.float64).reshape(-1) * (d.eps_hi - d.eps_lo)

    def set_density(self, rho: torch.Tensor):
        if rho.numel() != self.design.n_voxels:
            raise ValueError(f"rho has {rho.numel()} values, the design region {self.design.n_voxels}")
        self.set_permittivity(self.eps(rho.to(self._dev)))

    def set_permittivity(self, eps: torch.Tensor):
        """Permittivity per design voxel (C order over the box), set at the voxel's E_x, E_y, E_z nodes."""
        eps = eps.to(device=self._dev, dtype=torch.float64).reshape(-1)
        if eps.numel() != self.design.n_voxels:
            raise ValueError(f"eps has {eps.numel()} values, the design region {self.design.n_voxels}")
        if float(eps.min()) <= 0:
            raise ValueError("the design permittivity must be positive")
        cb = (self.fwd.cdx / eps)[self._vox].to(torch.float32)
        self.fwd._mat[self._nodes] = cb
        self._adj_mat[self._nodes] = cb
        self.cb_design.copy_(cb)
        self._cb_e.copy_(self._adj_mat[self._e_rows.entry_target])

    # --- runs -----------------------------------------------------------------------------------------
    def _zero(self, eng: Engine):
        for t in eng.state_tensors():
            t.zero_()

    def forward(self, rho: torch.Tensor):
        """Forward run for density rho: the monitor's (acc_e, acc_h) after n_steps (copies)."""
        self.set_density(rho)
        return self._forward_run()

    def forward_permittivity(self, eps: torch.Tensor):
        """Forward run with the design voxels' permittivity given directly: the monitor's (acc_e, acc_h)."""
        self.set_permittivity(eps)
        return self._forward_run()

    def _forward_run(self):
        self._zero(self.fwd)
        self.runs += 1
        fwd = self.fwd
        self.base.fill_(0 if self.store_all else -(1 << 62))
        self.checkpoints = []
        if self._run is None:
            self._run, _ = fwd.make_runner(self.use_graphs)
        for step in range(0, self.n_steps, self.stride):
            if not self.store_all and step % self.K == 0:
                self.checkpoints.append([t.clone() for t in fwd.state_tensors()])
            self._run()
        m = self.monitor
        acc_h = m.acc_h.clone() if m.acc_h is not None else torch.zeros(0, dtype=torch.complex64, device=self._dev)
        return m.acc_e.clone(), acc_h

    def begin_backward(self, grad_e: torch.Tensor | None, grad_h: torch.Tensor | None):
        """Zero the adjoint state, load the monitor's adjoint source weights and set the adjoint's initial condition
        (the sample of E^{n_steps})."""
        adj = self.adj
        self._zero(adj)
        self.g.zero_()
        self.W_E.copy_(torch.conj(grad_e) if grad_e is not None else torch.zeros_like(self.W_E))
        if self.W_H is not None:
            self.W_H.copy_(torch.conj(grad_h) if grad_h is not None and grad_h.numel() else torch.zeros_like(self.W_H))
        t = self.n_steps * self.dt
        ph = torch.exp(1j * self._omega * t).to(torch.complex128)
        src = torch.real(self.W_E.to(torch.complex128) * ph[:, None]).sum(0)
        per_node = self._e_rows.row_sum(src, self._cb_e.to(torch.float64))
        adj.E_flat[self._e_nodes] += per_node.to(torch.float32)

    def restore_checkpoint(self, seg: int):
        """The forward state at the start of checkpoint interval seg, with the tape rows of that interval next."""
        for tt, snap in zip(self.fwd.state_tensors(), self.checkpoints[seg]):
            tt.copy_(snap)
        self.base.fill_(seg * self.K)

    def interval(self, seg: int) -> int:
        return min(self.K, self.n_steps - seg * self.K)

    def finish_backward_permittivity(self) -> torch.Tensor:
        """dL/deps per design voxel (float64) from the accumulated dL/dcb: cb = C / eps, dcb/deps = -cb^2 / C."""
        cb = self.cb_design.to(torch.float64)
        return (self.g * (-(cb * cb) / self.fwd.cdx)).view(3, self.design.n_voxels).sum(0)

    def backward_permittivity(self, grad_e: torch.Tensor | None, grad_h: torch.Tensor | None) -> torch.Tensor:
        """dL/deps per design voxel for the last forward run (forward or forward_permittivity)."""
        self._adjoint_pass(grad_e, grad_h)
        return self.finish_backward_permittivity()

    def finish_backward(self) -> torch.Tensor:
        """dL/drho (float64) from the accumulated dL/dcb of the design nodes."""
        d = self.design
        cb = self.cb_design.to(torch.float64)
        # cb = C / eps  ->  dcb/deps = -cb^2 / C ; deps/drho = eps_hi - eps_lo
        dcb_drho = -(cb * cb) / self.fwd.cdx * (d.eps_hi - d.eps_lo)
        # nodes are component-major over the voxels (DesignRegion.nodes): a fixed-order sum, no atomics
        return (self.g * dcb_drho).view(3, d.n_voxels).sum(0)

    def backward(self, grad_e: torch.Tensor | None, grad_h: torch.Tensor | None) -> torch.Tensor:
        """dL/drho (float64, the shape of rho's flattening) for the objective whose gradients with respect to the
        monitor's accumulators are grad_e, grad_h (torch's convention for complex tensors)."""
        self._adjoint_pass(grad_e, grad_h)
        return self.finish_backward()

    def _adjoint_pass(self, grad_e, grad_h):
        self.begin_backward(grad_e, grad_h)
        if self.store_all:
            self.base.fill_(0)
            self.adj.steps(self.n_steps)
        else:
            for seg in reversed(range(len(self.checkpoints))):
                self.restore_checkpoint(seg)
                self.fwd.steps(self.interval(seg))
                self.adj.steps(self.interval(seg))

    def release_graphs(self):
        """Drop the forward engine's captured chunk graph (before the engine joins a batch)."""
        if self._run is not None:
            self.fwd.release()
            self._run = None

    def close(self):
        self.release_graphs()
        self.fwd.close()
        self.adj.close()


class BatchedAdjoint:
    """Forward and adjoint runs of several AdjointProblems (multi-instance inverse design: uneven grids, each with
    its own design region, sources, dispersive groups and monitor) stepped together: the forward engines as one
    Batch, the adjoint engines as another, so every half-step is one launch per kernel configuration over all
    instances' tiles (and one HIP graph per forward DFT chunk). With checkpoints the adjoint pass recomputes each
    interval of every instance in one batched forward run right before reversing it in one batched adjoint run.
    Each instance's result equals its run alone bit for bit (the batched kernels run the same tile code)."""

    def __init__(self, problems, use_graphs: bool = True, schedule=None):
        from .batch import Batch
        self.problems = list(problems)
        if not self.problems:
            raise ValueError("a batched adjoint needs at least one problem")
        p0 = self.problems[0]
        for p in self.problems:
            if (p.stride, p.K, p.n_steps, p.store_all) != (p0.stride, p0.K, p0.n_steps, p0.store_all):
                raise ValueError("batched adjoint problems share the DFT stride, checkpoint interval, step count "
                                 "and tape mode")
        self.stride, self.K, self.n_steps, self.store_all = p0.stride, p0.K, p0.n_steps, p0.store_all
        self.use_graphs = use_graphs
        # one row per lane and 64-cell rows as the members' own tiling (engines with up to 4 poles per node need
        # one row per lane to keep registers; the adjoint kernels take 64-cell rows), with the z chunk that fills
        # the GPU over all members together, one block per tile
        ry, waves, _, zc = filling_tiling([p.fwd.shape for p in self.problems])
        sched = schedule or (ry, waves, 64, zc, 0)
        self.fwd = Batch([p.fwd for p in self.problems], schedule=sched)
        self.adj = Batch([p.adj for p in self.problems], schedule=sched)
        self._run = None
        self.runs = 0

    def forward(self, rhos) -> list:
        """[(acc_e, acc_h)] of every problem after a batched forward run with the given densities."""
        for p, rho in zip(self.problems, rhos):
            p.set_density(rho)
            p._zero(p.fwd)
            p.runs += 1
            p.base.fill_(0 if p.store_all else -(1 << 62))
            p.checkpoints = []
        self.runs += 1
        if self._run is None:
            self._run, _ = self.fwd.make_runner(self.use_graphs)
        for step in range(0, self.n_steps, self.stride):
            if not self.store_all and step % self.K == 0:
                for p in self.problems:
                    p.checkpoints.append([t.clone() for t in p.fwd.state_tensors()])
            self._run()
        out = []
        for p in self.problems:
            m = p.monitor
            acc_h = m.acc_h.clone() if m.acc_h is not None else torch.zeros(0, dtype=torch.complex64, device=p._dev)
            out.append((m.acc_e.clone(), acc_h))
        return out

    def backward(self, grads) -> list:
        """[dL/drho] for the objectives whose accumulator gradients are grads = [(grad_e, grad_h)]."""
        for p, (ge, gh) in zip(self.problems, grads):
            p.begin_backward(ge, gh)
        if self.store_all:
            for p in self.problems:
                p.base.fill_(0)
            self.adj.steps(self.n_steps)
        else:
            for seg in reversed(range(len(self.problems[0].checkpoints))):
                for p in self.problems:
                    p.restore_checkpoint(seg)
                n_seg = self.problems[0].interval(seg)
                self.fwd.steps(n_seg)
                self.adj.steps(n_seg)
        return [p.finish_backward() for p in self.problems]

    def release(self):
        if self._run is not None:
            self.fwd.release()
            self._run = None

    def close(self):
        self.release()
        self.fwd.close()
        self.adj.close()


class _BatchedSpectra(torch.autograd.Function):
    @staticmethod
    def forward(ctx, batch, *rhos):
        ctx.batch = batch
        ctx.shapes = [r.shape for r in rhos]
        ctx.dtypes = [r.dtype for r in rhos]
        outs = batch.forward([r.detach() for r in rhos])
        ctx.run = batch.runs
        ctx.save_for_backward(*[r.detach().clone() for r in rhos])
        return tuple(t for pair in outs for t in pair)

    @staticmethod
    def backward(ctx, *grads):
        batch = ctx.batch
        if batch.runs != ctx.run:
            batch.forward(list(ctx.saved_tensors))
        pairs = [(grads[2 * q], grads[2 * q + 1]) for q in range(len(batch.problems))]
        gs = batch.backward(pairs)
        return (None, *[g.reshape(s).to(d) for g, s, d in zip(gs, ctx.shapes, ctx.dtypes)])


def batched_monitor_spectra(rhos, batch: BatchedAdjoint) -> list:
    """[(acc_e, acc_h)] of every problem of the batch after one batched forward run with densities rhos;
    differentiable in every rho through one batched adjoint run."""
    flat = _BatchedSpectra.apply(batch, *rhos)
    return [(flat[2 * q], flat[2 * q + 1]) for q in range(len(rhos))]


class _MonitorSpectra(torch.autograd.Function):
    @staticmethod
    def forward(ctx, rho, prob):
        ctx.prob = prob
        ctx.shape = rho.shape
        ctx.dtype = rho.dtype
        out = prob.forward(rho.detach())
        ctx.run = prob.runs
        ctx.save_for_backward(rho.detach().clone())
        return out

    @staticmethod
    def backward(ctx, grad_e, grad_h):
        prob = ctx.prob
        if prob.runs != ctx.run:
            # the problem ran another density since this output was made (two designs in one loss, a graph kept
            # across a line search): the tape, checkpoints and design cb are that run's, so redo this one first
            (rho,) = ctx.saved_tensors
            prob.forward(rho)
        g = prob.backward(grad_e, grad_h)
        return g.reshape(ctx.shape).to(ctx.dtype), None


def monitor_spectra(rho: torch.Tensor, prob: AdjointProblem):
    """(acc_e, acc_h) of the problem's DFT monitor after a forward run with density rho; differentiable in rho
    through the adjoint run."""
    return _MonitorSpectra.apply(rho, prob)
