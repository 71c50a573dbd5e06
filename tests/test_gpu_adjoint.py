"""The hand-written adjoint (transposed, time-reversed Yee + CPML update; lsf_set_adjoint) against autograd
through the float64 torch per-step restatement of the same update (oracle/torch_fdtd.py, itself pinned to the
FP64 oracle and its finite differences in test_adjoint_oracle.py).

Tolerance: relative L2 error of dL/drho over the design region < 1e-3 (the float32 GPU run against the float64
reference). Checkpoint recompute must reproduce the stored-tape gradient exactly.
"""

import os

import numpy as np
import pytest
import torch

from lightsmith_amd.fdtd.adjoint import monitor_spectra
from oracle.torch_fdtd import TorchRun
from tests import cases

pytestmark = pytest.mark.gpu

GRAD_TOL = 1e-3


def _hip_grad(case, gpu, **kw):
    prob = cases.adjoint_problem(case, gpu, **kw)
    rho = torch.tensor(case["rho"], dtype=torch.float64, device=gpu, requires_grad=True)
    acc_e, acc_h = monitor_spectra(rho, prob)
    loss = cases.adjoint_objective(acc_e, acc_h, case)
    loss.backward()
    out = (rho.grad.detach().cpu().numpy(), float(loss.detach()), acc_e.detach().cpu().numpy(),
           acc_h.detach().cpu().numpy(), prob)
    return out


def _torch_grad(case):
    run = TorchRun(cases.adjoint_oracle(case, case["rho"]), case["lo"], case["hi"], case["eps_lo"], case["eps_hi"])
    rho = torch.as_tensor(case["rho"]).requires_grad_(True)
    te, th = run.run(rho, case["steps"])
    loss = cases.adjoint_objective(te, th, case)
    loss.backward()
    return rho.grad.numpy(), float(loss.detach()), te.detach().numpy(), th.detach().numpy()


def test_adjoint_gradient_matches_autograd(gpu):
    case = cases.adjoint_case()
    g, loss, e, h, prob = _hip_grad(case, gpu)
    g_ref, loss_ref, e_ref, h_ref = _torch_grad(case)
    assert prob.store_all
    err_e = cases.rel_l2(e, e_ref)
    err_g = cases.rel_l2(g, g_ref)
    print(f"loss {loss:.6e} (float64 {loss_ref:.6e}); spectra rel L2 {err_e:.2e}; dL/drho rel L2 {err_g:.2e} over "
          f"{g.size} voxels, max |g| {np.abs(g_ref).max():.3e}")
    assert err_e < 1e-5
    assert err_g < GRAD_TOL
    prob.close()


def test_checkpoint_recompute_reproduces_the_stored_tape(gpu):
    case = cases.adjoint_case(steps=200)
    g_all, *_, p1 = _hip_grad(case, gpu)
    g_ck, *_, p2 = _hip_grad(case, gpu, checkpoint_every=48, tape_budget=1)
    assert p1.store_all and not p2.store_all and len(p2.checkpoints) == 5
    print(f"stored tape vs checkpoint recompute: max |diff| {np.abs(g_all - g_ck).max():.3e}")
    assert np.array_equal(g_all, g_ck)
    p1.close()
    p2.close()


def test_adjoint_through_dispersive_nodes_matches_autograd(gpu):
    """A Drude + Lorentz D-form block between the design box and the monitor: the adjoint runs the transposed pole
    recursion (lsf_add_adjoint_ade). Gradient against float64 autograd through the torch restatement with the same
    pure group; checkpoint recompute reproduces the stored tape exactly (the pole state is part of a checkpoint)."""
    case = cases.adjoint_case(ade=True, steps=240)
    g, loss, e, h, prob = _hip_grad(case, gpu)
    g_ref, loss_ref, e_ref, _ = _torch_grad(case)
    assert prob.fwd.n_ade == len(case["ade_idx"]) and prob.store_all
    g_ck, *_, p2 = _hip_grad(case, gpu, checkpoint_every=48, tape_budget=1)
    err_e, err_g = cases.rel_l2(e, e_ref), cases.rel_l2(g, g_ref)
    print(f"ADE adjoint: {prob.fwd.n_ade} D-form nodes; spectra rel L2 {err_e:.2e}; dL/drho rel L2 {err_g:.2e}; "
          f"checkpointed max |diff| {np.abs(g - g_ck).max():.3e}")
    assert err_e < 1e-5 and err_g < GRAD_TOL
    assert np.array_equal(g, g_ck)
    prob.close()
    p2.close()


def test_flux_box_monitor_gradient_is_deterministic_and_matches_autograd(gpu):
    """A closed flux-box monitor: points on its edges and corners share E and H nodes, so the adjoint source
    sums several points into one node (fixed-order CSR rows). Two runs give bit-identical gradients, checkpoint
    recompute reproduces the stored tape exactly, and the gradient matches float64 autograd."""
    case = cases.adjoint_case(steps=200, monitor="box")
    e_nodes = case["e_mon"]
    assert len(np.unique(e_nodes)) < len(e_nodes)  # shared nodes exist
    g1, _, e1, _, p1 = _hip_grad(case, gpu)
    g2, *_, p2 = _hip_grad(case, gpu)
    g3, *_, p3 = _hip_grad(case, gpu, checkpoint_every=48, tape_budget=1)
    g_ref, _, e_ref, _ = _torch_grad(case)
    err = cases.rel_l2(g1, g_ref)
    print(f"flux-box monitor: dL/drho rel L2 {err:.2e}; spectra {cases.rel_l2(e1, e_ref):.2e}")
    assert np.array_equal(g1, g2) and np.array_equal(g1, g3)
    assert err < GRAD_TOL
    for p in (p1, p2, p3):
        p.close()


def test_interleaved_forwards_before_backward(gpu):
    """Two densities run through one problem before a single backward of a loss over both outputs: each output's
    gradient belongs to its own density (the problem re-runs a forward whose tape was overwritten)."""
    case = cases.adjoint_case(steps=160)
    prob = cases.adjoint_problem(case, gpu)
    rho_a = torch.tensor(case["rho"], dtype=torch.float64, device=gpu, requires_grad=True)
    rho_b = torch.tensor(1.0 - case["rho"], dtype=torch.float64, device=gpu, requires_grad=True)
    ea, ha = monitor_spectra(rho_a, prob)
    eb, hb = monitor_spectra(rho_b, prob)
    (cases.adjoint_objective(ea, ha, case) + 2.0 * cases.adjoint_objective(eb, hb, case)).backward()
    ga, gb = rho_a.grad.cpu().numpy(), rho_b.grad.cpu().numpy()
    ref_a, *_ = _torch_grad(case)
    ref_b, *_ = _torch_grad(dict(case, rho=1.0 - case["rho"]))
    print(f"interleaved: rel L2 {cases.rel_l2(ga, ref_a):.2e}, {cases.rel_l2(gb, 2.0 * ref_b):.2e}")
    assert cases.rel_l2(ga, ref_a) < GRAD_TOL and cases.rel_l2(gb, 2.0 * ref_b) < GRAD_TOL
    prob.close()


def test_adjoint_refuses_what_it_does_not_transpose(gpu):
    case = cases.adjoint_case()
    with pytest.raises(ValueError):
        cases.adjoint_problem(dict(case, lo=(2, 9, 8)), gpu)  # design region inside the CPML


def test_full_scene_gradient_matches_float64_central_differences(gpu):
    """The 256x192x96 adjoint scene (3000 steps): dL/drho on seeded voxels (seed 5) against central differences of
    the float64 torch restatement run on the GPU (h = 1e-3): the first 6 of the 32 seeded voxels (two float64 runs
    of about 4 s each per voxel), all 32 with LSF_ADJOINT_FULL=1."""
    from lightsmith_amd.fdtd import workloads
    scene = workloads.adjoint_scene()
    prob = workloads.adjoint_problem(scene, gpu)
    rho = torch.tensor(scene["rho"], dtype=torch.float64, device=gpu, requires_grad=True)
    acc_e, _ = monitor_spectra(rho, prob)
    workloads.mode_power(acc_e, scene["w"]).backward()
    g = rho.grad.cpu().numpy()
    prob.close()
    case = dict(scene, steps=scene["n_steps"], kappa_max=1.0, alpha_max=0.05 * float(scene["omega"][0]))
    o = cases.adjoint_oracle(case, scene["rho"])
    run = TorchRun(o, scene["lo"], scene["hi"], scene["eps_lo"], scene["eps_hi"], device=gpu)
    rng = np.random.default_rng(5)
    picks = [tuple(int(rng.integers(0, s)) for s in scene["rho"].shape) for _ in range(32)]
    picks = picks if os.environ.get("LSF_ADJOINT_FULL") else picks[:6]
    h = 1e-3
    fd, ad = [], []
    with torch.no_grad():
        for q, v in enumerate(picks):
            vals = []
            for sgn in (1, -1):
                r = scene["rho"].copy()
                r[v] += sgn * h
                e, _ = run.run(torch.as_tensor(r), scene["n_steps"])
                vals.append(float(workloads.mode_power(e.cpu().numpy(), scene["w"])))
            fd.append((vals[0] - vals[1]) / (2 * h))
            ad.append(g[v])
            print(f"voxel {q} {v}: adjoint {ad[-1]:.6e}  central difference {fd[-1]:.6e}", flush=True)
    fd, ad = np.array(fd), np.array(ad)
    err = np.linalg.norm(ad - fd) / np.linalg.norm(fd)
    print(f"{len(picks)} voxels: relative L2 error {err:.3e}; max per-voxel relative error "
          f"{np.max(np.abs(ad - fd) / np.abs(fd)):.3e}")
    assert err < GRAD_TOL
