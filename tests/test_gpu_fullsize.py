"""Size-independent properties at the full benchmark sizes, where the FP64 oracle would take hours.

* 320^3 throughput scene (16-cell CPML, 200 spheres, 7.16 M D-form 3-pole nodes, dipole, 16-frequency DFT plane):
  two engines with different launch tilings (z chunk, wavefronts per block) end bit-identical after 100 steps, the
  arithmetic of every cell being the same whatever tile computes it; the single-pass step (another kernel, double
  buffers) agrees with the two-kernel step to rounding; the energy proxy and the DFT plane are finite and non-zero.
* 64-run oblique sweep (91.8 M complex and real cells, 48-96 x 48-96 x 256): the autotuner's schedule and a batch
  of all 64 runs end bit-identical (every state tensor) to the same runs stepped one after another.
The small-grid versions of both checks, against the oracle, are in test_gpu_parity.py and test_gpu_batch.py.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

from lightsmith_amd.fdtd import workloads
from lightsmith_amd.fdtd.batch import Batch

pytestmark = pytest.mark.gpu


def _state_equal(a, b) -> bool:
    ta, tb = a.state_tensors(), b.state_tensors()
    return len(ta) == len(tb) and all(x.shape == y.shape and torch.equal(x, y) for x, y in zip(ta, tb))


def _rel(x: torch.Tensor, y: torch.Tensor) -> float:
    return float((x.double() - y.double()).norm() / y.double().norm())


def test_throughput_320_tilings_bit_identical_and_single_pass_to_rounding(gpu):
    scene = workloads.throughput_scene(n=320)
    steps = 2 * scene["stride"]  # two DFT chunks: 100 steps
    a = workloads.throughput_engine(scene, gpu, max_steps=steps + 10)
    b = workloads.throughput_engine(scene, gpu, max_steps=steps + 10, tiling=(1, 8, 64, 16))
    assert a.tiling != b.tiling
    for e in (a, b):
        run, used = e.make_runner(True)
        assert used
        for _ in range(steps // e.stride):
            run()
        e.release()
    torch.cuda.synchronize()
    same = _state_equal(a, b)
    en = a.energy.double().cpu().numpy()
    ea, _ = a.monitors[0].results()
    print(f"320^3, {steps} steps: tilings {a.tiling} / {b.tiling} bit-identical: {same}; energy {en}")
    assert same
    assert np.all(np.isfinite(en)) and en[0] > 0
    assert np.all(np.isfinite(ea)) and np.abs(ea).max() > 0
    del b
    torch.cuda.empty_cache()
    c = workloads.throughput_engine(scene, gpu, max_steps=steps + 10, fused=True)
    assert c._two_sets
    for _ in range(steps // c.stride):
        c.chunk()
    torch.cuda.synchronize()
    re, rh = _rel(c.E_flat, a.E_flat), _rel(c.H_flat, a.H_flat)
    ec, _ = c.monitors[0].results()
    rd = float(np.linalg.norm(ec - ea) / np.linalg.norm(ea))
    print(f"single pass vs two kernels after {steps} steps: E {re:.1e}, H {rh:.1e}, DFT {rd:.1e}")
    assert re < 2e-5 and rh < 2e-5 and rd < 2e-5


def test_oblique_sweep_64_runs_batched_bit_identical_to_sequential(gpu):
    runs = workloads.oblique_sweep_scene()
    stride = workloads.common_stride(runs)
    steps = 2 * stride
    alone = [workloads.oblique_engine(r, gpu, max_steps=steps + 4 * stride, stride=stride) for r in runs]
    for e in alone:
        run, _ = e.make_runner(True)
        for _ in range(steps // stride):
            run()
        e.release()
    members = [workloads.oblique_engine(r, gpu, max_steps=steps + 4 * stride, stride=stride) for r in runs]
    b = Batch(members, schedule=(1, 4, 32, 16, 0))  # the schedule the autotuner picks for this sweep on MI355X
    run, graphs = b.make_runner(True)
    assert graphs
    for _ in range(steps // stride):
        run()
    b.release()
    torch.cuda.synchronize()
    cells = sum(e.N for e in members)
    diff = [q for q, (x, y) in enumerate(zip(alone, members)) if not _state_equal(x, y)]
    print(f"oblique sweep: {len(runs)} runs, {cells / 1e6:.1f} M cells, {steps} steps; members differing: {diff}")
    assert len(runs) == 64 and cells > 90e6
    assert not diff
    b.close()
