"""lightsmith's differentiable-backend protocol for inverse design (inverse/__init__.py:14-40, 84-99 and
topopt.py:85-89 of the reference) over the FDTD adjoint of this package.

The reference's topology optimiser drives any backend with ``differentiable = True`` through three classmethods:

    design_grid(problem) -> grid                (grid.x, grid.y, grid.z in nm, grid.dl, grid.shape)
    base_permittivity(problem) -> (n_wl, *grid.shape) complex
    differentiable_solve(problem, eps_grid_override, device, sources=None, with_fields=True, **context)
        -> solution with torch tensors .s (n_wl, n_out, n_src) complex, .outputs, .inputs, .fields, .grid

and an optional ``prepare(problem) -> dict`` whose result comes back as ``**context`` on every call. Here the
problem is an ``FDTDDesignProblem``: a 3D FDTD scene (grid, materials, CPML, a mode-injection source, an output
monitor with mode-overlap weights) and the box of cells whose permittivity the optimiser sets. The design grid is
that box; ``differentiable_solve`` runs one forward FDTD on the GPU and, when the loss is differentiated, one adjoint
run (time-reversed, transposed Yee + CPML + ADE update with checkpoint recompute), so gradients with respect to every
cell of the box are exact for the discretised problem. The single S-parameter is the mode overlap at the output
divided by the source current's spectrum at each wavelength (one output 'o2', one input 'o1').

Port-mode (FDE) source tables, S-parameter matrices over several ports and 2D devices stay with the reference's
FDFD backend; this adapter is what ``inverse.optimize`` needs to run an FDTD adjoint on such a scene.
"""

from __future__ import annotations

from dataclasses import dataclass, field
from typing import Any, ClassVar

import numpy as np
import torch


@dataclass
class FDTDDesignProblem:
    """A scene dict in the form of ``fdtd.workloads.adjoint_scene`` (shape, dx, dt, inv_eps, pml, lo/hi design box,
    eps_lo/eps_hi bounds, source and monitor index tables, overlap weights w, omega, n_steps, stride; optionally
    dispersive ``groups`` outside the box), and the checkpoint settings of its adjoint."""

    scene: dict
    checkpoint_every: int = 100
    tape_budget: int = 4 << 30

    @property
    def wavelengths_nm(self) -> np.ndarray:
        return 2 * np.pi / np.asarray(self.scene["omega"], dtype=float)


@dataclass
class DesignGrid:
    """Node coordinates (nm) of the design box's cells, cell size, shape (the protocol's ``grid``)."""

    x: np.ndarray
    y: np.ndarray
    z: np.ndarray
    dl: float
    shape: tuple


@dataclass
class DifferentiableSolution:
    s: torch.Tensor
    outputs: list = field(default_factory=lambda: ["o2"])
    inputs: list = field(default_factory=lambda: ["o1"])
    fields: Any = None
    grid: DesignGrid | None = None


class _EpsSpectra(torch.autograd.Function):
    """acc_e of the problem's monitor as a function of the box permittivity; backward by one adjoint run."""

    @staticmethod
    def forward(ctx, eps, prob):
        ctx.prob = prob
        acc_e, _ = prob.forward_permittivity(eps.detach().reshape(-1))
        ctx.run = prob.runs
        ctx.shape, ctx.dtype = eps.shape, eps.dtype
        ctx.save_for_backward(eps.detach().clone())
        return acc_e

    @staticmethod
    def backward(ctx, grad_e):
        prob = ctx.prob
        if prob.runs != ctx.run:
            (eps,) = ctx.saved_tensors
            prob.forward_permittivity(eps.reshape(-1))
        g = prob.backward_permittivity(grad_e, None)
        return g.reshape(ctx.shape).to(ctx.dtype), None


class DifferentiableFDTD:
    """The protocol's classmethods for ``FDTDDesignProblem`` (inverse/__init__.py:84-99); ``check_backend``
    (topopt.py:85-89) accepts it."""

    name: ClassVar[str] = "lightsmith-fdtd"
    method: ClassVar[str] = "fdtd"
    differentiable: ClassVar[bool] = True

    @classmethod
    def design_grid(cls, problem: FDTDDesignProblem) -> DesignGrid:
        sc = problem.scene
        dx = float(sc["dx"])
        lo, hi = sc["lo"], sc["hi"]
        ax = [(np.arange(lo[a], hi[a]) - sc["shape"][a] / 2) * dx for a in range(3)]
        return DesignGrid(ax[0], ax[1], ax[2], dx, tuple(int(h - l) for l, h in zip(lo, hi)))

    @classmethod
    def base_permittivity(cls, problem: FDTDDesignProblem) -> np.ndarray:
        """(1, *grid.shape): the scene's own permittivity over the box (of its E_x nodes; the box is staircased)."""
        sc = problem.scene
        box = tuple(slice(l, h) for l, h in zip(sc["lo"], sc["hi"]))
        return (1.0 / np.asarray(sc["inv_eps"])[0][box]).astype(complex)[None]

    @classmethod
    def prepare(cls, problem: FDTDDesignProblem, device="cuda:0") -> dict:
        """The forward and adjoint engines, built once and reused by every differentiable_solve."""
        from .fdtd import workloads
        prob = workloads.adjoint_problem(problem.scene, device, checkpoint_every=problem.checkpoint_every,
                                         tape_budget=problem.tape_budget)
        return {"prob": prob}

    @classmethod
    def differentiable_solve(cls, problem: FDTDDesignProblem, eps_grid_override=None, device="cuda:0", sources=None,
                             with_fields: bool = True, **context) -> DifferentiableSolution:
        if sources not in (None, ["o1"], ("o1",)):
            raise ValueError("the FDTD design problem has one source port, 'o1'")
        prob = context.get("prob") or cls.prepare(problem, device)["prob"]
        grid = cls.design_grid(problem)
        if eps_grid_override is None:
            eps_grid_override = torch.as_tensor(cls.base_permittivity(problem).real, device=prob._dev)
        eps = eps_grid_override
        if torch.is_complex(eps):
            if float(eps.imag.abs().max()) > 0:
                raise ValueError("the FDTD design region takes a real permittivity")
            eps = eps.real
        if eps.dim() == 4:
            if eps.shape[0] != 1:
                raise ValueError("the FDTD design problem runs one wavelength per solve")
            eps = eps[0]
        if tuple(eps.shape) != grid.shape:
            raise ValueError(f"eps_grid_override has shape {tuple(eps.shape)}, the design grid {grid.shape}")
        acc_e = _EpsSpectra.apply(eps.to(device=prob._dev, dtype=torch.float64), prob)
        w = torch.as_tensor(problem.scene["w"], device=prob._dev).to(acc_e.dtype)
        norm = torch.as_tensor(_source_spectrum(problem.scene), device=prob._dev).to(acc_e.dtype)
        s = ((w[None, :] * acc_e).sum(-1) / norm)[:, None, None]
        return DifferentiableSolution(s=s, fields=None, grid=grid)


def _source_spectrum(scene) -> np.ndarray:
    """The mode source's time series (sampled at the half steps it is injected on) transformed at the monitor's
    frequencies: the input amplitude the S-parameter is normalised by."""
    t = (np.arange(len(scene["t_src"])) + 0.5) * float(scene["dt"])
    omega = np.asarray(scene["omega"], dtype=float)
    return np.exp(1j * np.outer(omega, t)) @ np.asarray(scene["t_src"])[:, 0]


__all__ = ["DesignGrid", "DifferentiableFDTD", "DifferentiableSolution", "FDTDDesignProblem"]
