"""The torch per-step restatement (the autograd reference of the adjoint gradient) against the FP64 oracle:
same forward spectra, and autograd gradients equal to central finite differences of the oracle."""

import numpy as np
import pytest
import torch

from oracle.torch_fdtd import TorchRun
from tests import cases

SMALL = dict(shape=(22, 18, 16), pml=4, steps=64, stride=8, design=((8, 6, 5), (13, 12, 10)), wavelength=300.0)


def _oracle_loss(case, rho):
    eng = cases.adjoint_oracle(case, rho)
    for _ in range(case["steps"] // case["stride"]):
        eng.chunk()
    e, h = eng.monitors[0].acc_e, eng.monitors[0].acc_h
    return cases.adjoint_objective(e, h, case), e, h


@pytest.mark.parametrize("ade", [False, True])
def test_torch_run_reproduces_the_oracle_spectra(ade):
    case = cases.adjoint_case(**SMALL, ade=ade)
    rho = case["rho"]
    _, e, h = _oracle_loss(case, rho)
    run = TorchRun(cases.adjoint_oracle(case, rho), case["lo"], case["hi"], case["eps_lo"], case["eps_hi"])
    te, th = run.run(torch.as_tensor(rho), case["steps"])
    scale = np.abs(e).max()
    assert np.abs(te.numpy() - e).max() < 1e-12 * scale
    assert np.abs(th.numpy() - h).max() < 1e-12 * np.abs(h).max()


@pytest.mark.parametrize("ade", [False, True])
def test_torch_autograd_equals_oracle_finite_differences(ade):
    """With ade, a Drude + Lorentz D-form block sits between the design box and the monitor: the gradient flows
    through the pole recursion."""
    case = cases.adjoint_case(**SMALL, ade=ade)
    rho0 = case["rho"]
    run = TorchRun(cases.adjoint_oracle(case, rho0), case["lo"], case["hi"], case["eps_lo"], case["eps_hi"])
    rho = torch.as_tensor(rho0).requires_grad_(True)
    te, th = run.run(rho, case["steps"])
    cases.adjoint_objective(te, th, case).backward()
    g = rho.grad.numpy()
    rng = np.random.default_rng(7)
    picks = [tuple(rng.integers(0, s) for s in rho0.shape) for _ in range(4)]
    h = 1e-4
    for v in picks:
        rp, rm = rho0.copy(), rho0.copy()
        rp[v] += h
        rm[v] -= h
        fd = (_oracle_loss(case, rp)[0] - _oracle_loss(case, rm)[0]) / (2 * h)
        print(f"voxel {v}: autograd {g[v]:.6e}, oracle central difference {fd:.6e}")
        assert abs(g[v] - fd) <= 1e-6 * max(abs(fd), 1e-3 * np.abs(g).max())
