"""HIP engine vs the FP64 oracle on the same inputs (tests/cases.py).

Tolerance: relative L2 error of the float32 GPU fields against the float64 oracle after the stated number
of steps, FIELD_TOL = 1e-5, and DFT_TOL for the monitors' spectra. Each test
prints the measured error.
"""

import numpy as np
import pytest

from tests import cases

pytestmark = pytest.mark.gpu

FIELD_TOL = 1e-5
DFT_TOL = 1e-5


def _compare(o, g, label, mask=None, tol=FIELD_TOL):
    Eo, Ho = cases.fields(o)
    Eg, Hg = cases.fields(g)
    eE, eH = cases.rel_l2(Eg, Eo, mask), cases.rel_l2(Hg, Ho, mask)
    print(f"{label}: rel L2 E {eE:.2e}  H {eH:.2e}")
    assert eE < tol and eH < tol, (eE, eH)
    return eE, eH


def test_pec_box_random_fields(gpu):
    o = cases.make(cases.oracle_ns(), shape=(70, 33, 29), dx=1.0, dt=0.95 / np.sqrt(3),
                   inv_eps=np.full((3, 70, 33, 29), 1 / 2.25))
    g = cases.make(cases.product_ns(), gpu, shape=(70, 33, 29), dx=1.0, dt=0.95 / np.sqrt(3),
                   inv_eps=np.full((3, 70, 33, 29), 1 / 2.25))
    cases.random_fields(o, 1)
    cases.random_fields(g, 1)
    o.run(60)
    g.steps(60)
    _compare(o, g, "PEC box 60 steps")


@pytest.mark.parametrize("tiling", [(1, 4, 5), (2, 4, 16), (2, 8, 7), (1, 8, 3)])
def test_tilings_agree(gpu, tiling):
    shape = (67, 45, 23)
    eps = np.random.default_rng(3).uniform(1, 4, size=(3,) + shape)
    o = cases.make(cases.oracle_ns(), shape=shape, dx=1.0, dt=0.9 / np.sqrt(3), inv_eps=1 / eps,
                   pml=cases.oracle_ns().PMLSpec(cells=((5, 4), (3, 6), (4, 4)), alpha_max=0.02))
    ns = cases.product_ns()
    old = ns.Engine.tiling
    ns.Engine.tiling = tiling
    try:
        g = cases.make(ns, gpu, shape=shape, dx=1.0, dt=0.9 / np.sqrt(3), inv_eps=1 / eps,
                       pml=ns.PMLSpec(cells=((5, 4), (3, 6), (4, 4)), alpha_max=0.02))
    finally:
        ns.Engine.tiling = old
    cases.random_fields(o, 2)
    cases.random_fields(g, 2)
    o.run(40)
    g.steps(40)
    _compare(o, g, f"tiling {tiling}")


@pytest.mark.parametrize("material", ["indexed", "dense"])
def test_dispersive_cpml_dipole_box(gpu, material):
    steps = 300
    o = cases.dispersive_box(cases.oracle_ns(), n_steps_hint=steps)
    g = cases.dispersive_box(cases.product_ns(), gpu, n_steps_hint=steps, material=material)
    assert g.material_mode == material
    assert g.n_ade > 0
    for _ in range(steps // o.stride):
        o.chunk()
        g.chunk()
    _compare(o, g, f"dispersive box ({material}) {steps} steps")
    eo, ho = o.monitors[0].results()
    eg, hg = g.monitors[0].results()
    de, dh = cases.rel_l2(eg, eo), cases.rel_l2(hg, ho)
    print(f"DFT rel L2 E {de:.2e} H {dh:.2e}; energy {float(g.energy[0]):.6e} vs {o.energy[0]:.6e}")
    assert de < DFT_TOL and dh < DFT_TOL
    assert abs(float(g.energy[0]) / o.energy[0] - 1) < 1e-4


def test_cpml_kappa_alpha(gpu):
    steps = 200
    o = cases.dispersive_box(cases.oracle_ns(), n_steps_hint=steps, kappa_max=3.0, alpha=0.05, frac=0.0)
    g = cases.dispersive_box(cases.product_ns(), gpu, n_steps_hint=steps, kappa_max=3.0, alpha=0.05, frac=0.0)
    for _ in range(steps // o.stride):
        o.chunk()
        g.chunk()
    _compare(o, g, "CPML kappa/alpha")


def test_bloch_periodic_complex(gpu):
    shape = (20, 14, 40)
    dx = (3.0, 2.5, 2.0)
    dt = 0.9 / np.sqrt(sum(1 / d ** 2 for d in dx))
    kpar = (0.05, -0.03, 0.0)
    eps = np.random.default_rng(7).uniform(1, 6, size=(3,) + shape)
    kw = dict(shape=shape, dx=dx, dt=dt, inv_eps=1 / eps, periodic=(True, True, False), bloch=kpar)
    on, pn = cases.oracle_ns(), cases.product_ns()
    o = cases.make(on, pml=on.PMLSpec(cells=((0, 0), (0, 0), (6, 6))), **kw)
    g = cases.make(pn, gpu, pml=pn.PMLSpec(cells=((0, 0), (0, 0), (6, 6))), **kw)
    rng = np.random.default_rng(5)
    for c in range(3):
        for F, R in (("E", o.RE[c]), ("H", o.RH[c])):
            vals = rng.standard_normal(tuple(r.stop - r.start for r in R)) + 1j * rng.standard_normal(
                tuple(r.stop - r.start for r in R))
            getattr(o, F)[c][R] = vals
            import torch
            getattr(g, F)[c][R] = torch.as_tensor(vals, dtype=torch.complex64, device=g.E_flat.device)
    o.run(80)
    g.steps(80)
    _compare(o, g, "Bloch periodic complex", mask=cases.real_mask(shape, (True, True, False)))


def test_graph_replay_matches_eager(gpu):
    steps = 200
    a = cases.dispersive_box(cases.product_ns(), gpu, n_steps_hint=steps)
    b = cases.dispersive_box(cases.product_ns(), gpu, n_steps_hint=steps)
    run, used = b.make_runner(True)
    assert used
    for _ in range(steps // a.stride):
        a.chunk()
        run()
    b.release()
    Ea, Ha = cases.fields(a)
    Eb, Hb = cases.fields(b)
    assert np.array_equal(Ea, Eb) and np.array_equal(Ha, Hb)
    ea, _ = a.monitors[0].results()
    eb, _ = b.monitors[0].results()
    assert np.array_equal(ea, eb)
    assert int(a.n[0]) == int(b.n[0]) == steps
