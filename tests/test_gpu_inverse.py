"""The inverse-design adapter on the GPU: differentiable_solve's S-parameter and its gradient with respect to every
cell of the design box (one forward and one adjoint FDTD run) against float64 autograd through the per-step torch
restatement of the same scene (oracle/torch_fdtd.py), with dL/deps = dL/drho / (eps_hi - eps_lo)."""

import numpy as np
import pytest
import torch

from tests import cases

pytestmark = pytest.mark.gpu


def test_differentiable_solve_gradient_matches_autograd(gpu):
    from lightsmith_amd.fdtd import workloads
    from lightsmith_amd.inverse import DifferentiableFDTD, FDTDDesignProblem, _source_spectrum
    from oracle.torch_fdtd import TorchRun
    sc = workloads.adjoint_scene(shape=(96, 96, 64), design=(12, 10, 6), pml=10, n_steps=600, stride=25)
    prob = FDTDDesignProblem(sc)
    ctx = DifferentiableFDTD.prepare(prob, gpu)
    eps0 = sc["eps_lo"] + sc["rho"] * (sc["eps_hi"] - sc["eps_lo"])
    eps = torch.tensor(eps0, dtype=torch.float64, device=gpu, requires_grad=True)
    sol = DifferentiableFDTD.differentiable_solve(prob, eps[None], gpu, **ctx)
    assert sol.s.shape == (1, 1, 1) and sol.outputs == ["o2"] and sol.inputs == ["o1"]
    loss = sol.s.abs().pow(2).sum()
    loss.backward()
    g = eps.grad.cpu().numpy()
    # float64 autograd through the torch restatement, in rho, of the same loss
    case = dict(sc, steps=sc["n_steps"], kappa_max=1.0, alpha_max=0.05 * float(sc["omega"][0]))
    run = TorchRun(cases.adjoint_oracle(case, sc["rho"]), sc["lo"], sc["hi"], sc["eps_lo"], sc["eps_hi"], device=gpu)
    rho = torch.as_tensor(sc["rho"], device=gpu).requires_grad_(True)
    te, _ = run.run(rho, sc["n_steps"])
    w = torch.as_tensor(sc["w"], device=gpu).to(te.dtype)
    norm = torch.as_tensor(_source_spectrum(sc), device=gpu).to(te.dtype)
    s_ref = (w[None, :] * te).sum(-1) / norm
    s_ref.abs().pow(2).sum().backward()
    g_ref = rho.grad.cpu().numpy() / (sc["eps_hi"] - sc["eps_lo"])
    err_s = float(abs(sol.s.detach().cpu().numpy().ravel()[0] - s_ref.detach().cpu().numpy()[0]) /
                  abs(s_ref.detach().cpu().numpy()[0]))
    err_g = cases.rel_l2(g, g_ref)
    print(f"S-parameter rel err {err_s:.2e}; dL/deps over {g.size} cells rel L2 {err_g:.2e}")
    assert err_s < 1e-5 and err_g < 1e-3
    ctx["prob"].close()
