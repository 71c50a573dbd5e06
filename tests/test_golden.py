"""The committed golden fixtures (tests/golden/*.npz, written by tests/golden/make_golden.py from the FP64 oracle).

CPU: the oracle reproduces every fixture to 1e-12 relative L2 (a change to the oracle or to the seeded inputs of
tests/cases.py shows up here). GPU: the HIP engine, built from the same inputs, matches the stored fields and
spectra to the parity tolerance (1e-5 relative L2; energy 1e-4) without running the oracle.
"""

import numpy as np
import pytest

from tests import cases
from tests.golden import make_golden as mg

NAMES = list(mg.CASES)


def _load(name):
    with np.load(mg.path(name)) as z:
        return {k: z[k] for k in z.files}


def _check(out, ref, tol, m=None):
    errs = {}
    for k in ("E", "H"):
        errs[k] = cases.rel_l2(out[k], ref[k], m)
    for k in ("dft_e", "dft_h"):
        if k in ref and ref[k].size:
            errs[k] = cases.rel_l2(out[k], ref[k])
    for k, v in errs.items():
        assert v < tol, (k, v, tol)
    if "energy" in ref:
        e_tol = tol if tol < 1e-9 else 1e-4
        assert abs(out["energy"][0] - ref["energy"][0]) <= e_tol * abs(ref["energy"][0]), (out["energy"], ref["energy"])
    return errs


@pytest.mark.parametrize("name", NAMES)
def test_oracle_reproduces_golden(name):
    ref = _load(name)
    eng = mg.build(name, cases.oracle_ns())
    mg.run(eng, int(ref["steps"]))
    errs = _check(mg.outputs(eng), ref, 1e-12)
    print(name, {k: f"{v:.1e}" for k, v in errs.items()})


@pytest.mark.gpu
@pytest.mark.parametrize("name", NAMES)
def test_engine_matches_golden(gpu, name):
    import torch
    ref = _load(name)
    eng = mg.build(name, cases.product_ns(), gpu)
    mg.run(eng, int(ref["steps"]))
    torch.cuda.synchronize()
    out = mg.outputs(eng)
    errs = _check(out, ref, 1e-5, mg.mask(name, eng.shape))
    print(f"{name} vs golden:", {k: f"{v:.2e}" for k, v in errs.items()})
    eng.close()
