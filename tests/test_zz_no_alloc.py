"""lsf_step_noalloc (the stepping entry point that allocates nothing: its kernel arguments are staged in a
file-scope device array instead of a library allocation) against lsf_step through the C ABI: two calls in a row,
the second on a smaller engine with fewer steps, each bit-identical to lsf_step on a twin engine; an engine the
entry point cannot serve (a sub-engine) is refused with an error."""

import ctypes

import numpy as np
import pytest
import torch

from tests import cases

pytestmark = pytest.mark.gpu


def _bind(lib):
    fn = lib.lsf_step_noalloc
    fn.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]
    fn.restype = ctypes.c_int
    return fn


def _pair(gpu, shape, seed):
    ns = cases.product_ns()
    a = cases.dispersive_box(ns, gpu, shape=shape, seed=seed, n_steps_hint=200, kappa_max=2.0, alpha=0.02)
    b = cases.dispersive_box(ns, gpu, shape=shape, seed=seed, n_steps_hint=200, kappa_max=2.0, alpha=0.02)
    return a, b


def _state(eng):
    return [t.detach().cpu().clone() for t in eng.state_tensors()]


def test_noalloc_step_matches_lsf_step(gpu):
    from lightsmith_amd import _lib
    lib = _lib.load()
    noalloc = _bind(lib)
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    for shape, steps, seed in (((40, 36, 32), 37, 3), ((24, 22, 20), 11, 5)):
        ref, new = _pair(gpu, shape, seed)
        print(f"case {shape}, {steps} steps", flush=True)
        _lib.check(lib.lsf_step(ref._h, steps, stream), "lsf_step")
        rc = noalloc(new._h, steps, stream)
        _lib.check(rc, "lsf_step_noalloc")
        torch.cuda.synchronize()
        for x, y in zip(_state(ref), _state(new)):
            assert torch.equal(x, y)
        assert float(torch.linalg.norm(ref.E_flat.abs())) > 0
        ref.close()
        new.close()


def test_noalloc_step_refuses_sub_engines(gpu):
    """A Bloch plane source steps a column sub-engine inside the step: its arguments would need more scratch."""
    from lightsmith_amd import _lib
    from lightsmith_amd.fdtd import Engine, PMLSpec
    from lightsmith_amd.fdtd.sources import BlochPlaneSource
    lib = _lib.load()
    noalloc = _bind(lib)
    shape, sp = (11, 9, 40), (20.0, 24.0, 15.0)
    dt = 0.9 / np.sqrt(sum(1 / d ** 2 for d in sp))
    pml = PMLSpec(cells=((0, 0), (0, 0), (8, 8)))
    src = BlochPlaneSource(shape=shape, spacing=sp, dt=dt, k_src=20, k_launch=19, kpar=(0.004, 0.002), pol="p",
                           eps_b=1.0, pml=pml, omega_c=2 * np.pi / 600, omega_half=0.3 * 2 * np.pi / 600, max_steps=50)
    eng = Engine(gpu, shape=shape, dx=sp, dt=dt, single=True, inv_eps=np.ones((3,) + shape), pml=pml, sources=[src],
                 periodic=(True, True, False), bloch=(0.004, 0.002, 0.0))
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    assert noalloc(eng._h, 1, stream) != 0
    assert b"sub-engines" in lib.lsf_last_error()
    eng.close()
