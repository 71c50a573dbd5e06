"""The C restatement of the oracle step (oracle/cyee.c) against the NumPy oracle on the same configuration: fields
bit-identical (the same float64 operations in the same order, no contraction), DFT spectra and energy to 1e-12
(sin/cos and the energy sum may round differently)."""

import os
import subprocess

import numpy as np
import pytest

from tests import cases

ORACLE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle")


@pytest.fixture(scope="module")
def cyee():
    if not os.path.exists(os.path.join(ORACLE, "_cyee.so")):
        subprocess.run(["make", "-s", "-C", ORACLE], check=True)
    from oracle import cyee
    cyee.load()
    return cyee


@pytest.mark.parametrize("threads", ["1", "4"])
def test_c_step_equals_the_numpy_oracle(cyee, threads, monkeypatch):
    monkeypatch.setenv("OMP_NUM_THREADS", threads)
    ns = cases.oracle_ns()
    kw = dict(shape=(26, 22, 20), pml_cells=4, n_steps_hint=120, kappa_max=2.0, alpha=0.02, frac=0.15, seed=5)
    a = cases.dispersive_box(ns, **kw)
    b = cases.dispersive_box(ns, **kw)
    cases.random_fields(a, 3)
    cases.random_fields(b, 3)
    c = cyee.CYee(b)
    for _ in range(12):
        a.chunk()
    c.run(120)
    assert c.n == a.n == 120
    assert np.array_equal(a.E, c.E) and np.array_equal(a.H, c.H)
    for (_, _, _, _, _, _, _, pa), pc in zip(a.pml_E + a.pml_H, c.psi_e + c.psi_h):
        assert np.array_equal(pa, pc)
    ea, ha = a.monitors[0].results()
    ec, hc = c.results(0)
    assert cases.rel_l2(ec, ea) < 1e-12 and cases.rel_l2(hc, ha) < 1e-12
    assert abs(c.energy[0] / a.energy[0] - 1) < 1e-12


def test_c_step_refuses_what_it_does_not_restate(cyee):
    ns = cases.oracle_ns()
    shape = (10, 8, 12)
    eng = cases.make(ns, shape=shape, dx=(1.0, 1.0, 2.0), dt=0.3, inv_eps=np.ones((3,) + shape))
    with pytest.raises(ValueError):
        cyee.CYee(eng)  # non-uniform spacing
