"""Writes the golden fixtures of this directory from the FP64 oracle (``python -m tests.golden.make_golden``).

The reference ships no stored vectors for FDTD stepping and cannot run here, so these fixtures are outputs of the
oracle on seeded inputs (tests/cases.py), taken after the oracle was pinned by the closed forms and invariants of
tests/test_oracle_pins.py. They freeze those results: tests/test_golden.py checks that the oracle still reproduces
them (CPU) and that the HIP engine matches them (GPU) without running the oracle.

Each case file holds the inputs that are not code (step count, seeds and sizes as scalars) and the outputs:
E and H in the reference layout (3, N0, N1, N2) and, where the case has one, the DFT monitor's spectra and energy.
"""

from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

from tests import cases  # noqa: E402

# name -> (builder kwargs, steps); every builder takes the namespace first and a device for the HIP engine
CASES = {
    "dispersive_cpml_box": dict(steps=100),
    "bloch_complex": dict(steps=60),
    "tfsf_oblique": dict(steps=150),
}


def build(name, ns, device=None):
    """The engine of case `name` in namespace ns with its initial state set."""
    if name == "dispersive_cpml_box":
        return cases.dispersive_box(ns, device, shape=(24, 22, 20), pml_cells=4, n_steps_hint=100, kappa_max=2.0,
                                    alpha=0.02, frac=0.15, seed=11)
    if name == "bloch_complex":
        shape = (12, 10, 24)
        dx = (3.0, 2.5, 2.0)
        dt = 0.9 / np.sqrt(sum(1 / d ** 2 for d in dx))
        eps = np.random.default_rng(17).uniform(1, 6, size=(3,) + shape)
        eng = cases.make(ns, device, shape=shape, dx=dx, dt=dt, inv_eps=1 / eps, periodic=(True, True, False),
                         bloch=(0.07, -0.04, 0.0), pml=ns.PMLSpec(cells=((0, 0), (0, 0), (5, 5))))
        rng = np.random.default_rng(19)
        for c in range(3):
            for F, R in ((eng.E, eng.RE[c]), (eng.H, eng.RH[c])):
                size = tuple(r.stop - r.start for r in R)
                vals = rng.standard_normal(size) + 1j * rng.standard_normal(size)
                if hasattr(F, "cpu"):
                    import torch
                    F[c][tuple(R)] = torch.as_tensor(vals, dtype=F.dtype, device=F.device)
                else:
                    F[c][tuple(R)] = vals
        return eng
    if name == "tfsf_oblique":
        khat = np.array([1.0, 2.0, 3.0])
        k = khat / np.linalg.norm(khat)
        e = np.cross(k, [0.0, 1.0, 0.0])
        e /= np.linalg.norm(e)
        return cases.tfsf_box(ns, device, khat=tuple(khat), ehat=e, steps=150)
    raise KeyError(name)


def mask(name, shape):
    """Cells compared: the HIP engine does not materialise periodic ghost planes."""
    return cases.real_mask(shape, (True, True, False)) if name == "bloch_complex" else None


def run(eng, steps):
    if eng.monitors:
        for _ in range(steps // eng.stride):
            eng.chunk()
    else:
        eng.steps(steps) if hasattr(eng, "steps") else eng.run(steps)


def outputs(eng) -> dict:
    E, H = cases.fields(eng)
    out = {"E": E, "H": H}
    if eng.monitors:
        ae, ah = eng.monitors[0].results()
        out["dft_e"] = np.asarray(ae, complex)
        out["dft_h"] = np.asarray(ah, complex) if ah is not None else np.zeros(0, complex)
        en = getattr(eng, "energy", None)
        if en is not None:
            out["energy"] = np.asarray([float(en[0])])
    return out


def path(name) -> str:
    return os.path.join(HERE, f"{name}.npz")


def main():
    ns = cases.oracle_ns()
    for name, spec in CASES.items():
        eng = build(name, ns)
        run(eng, spec["steps"])
        out = outputs(eng)
        real = not (np.iscomplexobj(out["E"]) and np.abs(out["E"].imag).max() > 0)
        store = {k: (v.real if real and k in ("E", "H") else v) for k, v in out.items()}
        np.savez_compressed(path(name), steps=np.int64(spec["steps"]), **store)
        print(name, {k: v.shape for k, v in store.items()}, os.path.getsize(path(name)))


if __name__ == "__main__":
    main()
