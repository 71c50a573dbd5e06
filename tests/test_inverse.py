"""The inverse-design adapter's protocol surface on the CPU (inverse/__init__.py:84-99, topopt.py:85-89 of the
reference): differentiable, the three classmethods, a design grid and base permittivity of the design box."""

import numpy as np

from lightsmith_amd.fdtd import workloads
from lightsmith_amd.inverse import DifferentiableFDTD, FDTDDesignProblem


def _scene():
    return workloads.adjoint_scene(shape=(96, 96, 64), design=(12, 10, 6), pml=10, n_steps=600, stride=25)


def test_protocol_surface_passes_the_reference_backend_check():
    B = DifferentiableFDTD
    missing = [m for m in ("design_grid", "base_permittivity", "differentiable_solve") if not hasattr(B, m)]
    assert getattr(B, "differentiable", False) and not missing  # the test of check_backend (topopt.py:85-89)
    assert callable(B.prepare)


def test_design_grid_and_base_permittivity_cover_the_design_box():
    sc = _scene()
    prob = FDTDDesignProblem(sc)
    g = DifferentiableFDTD.design_grid(prob)
    assert g.shape == (12, 10, 6) and g.dl == sc["dx"]
    assert len(g.x) == 12 and len(g.y) == 10 and len(g.z) == 6
    assert np.allclose(np.diff(g.x), sc["dx"])
    base = DifferentiableFDTD.base_permittivity(prob)
    assert base.shape == (1, 12, 10, 6) and np.iscomplexobj(base)
    # the box sits in the waveguide core of the scene
    assert np.allclose(base.real, sc["eps_hi"]) or np.all(base.real >= sc["eps_lo"] - 1e-12)
    assert np.allclose(prob.wavelengths_nm, [1550.0])
