"""Batches (lsf_batch_*) against the same engines stepped one at a time.

A batch only changes how work is launched, not the arithmetic of any cell, source point or monitor: every
state tensor of every member (fields, psi, pole state, source lines, sub-engines, DFT accumulators, energy,
counters) must be bit-identical to a run of that member alone, for any schedule. The single-engine path is
itself checked against the FP64 oracle (test_gpu_parity.py); one member here is also checked against the
oracle directly.
"""

import numpy as np
import pytest
import torch

from lightsmith_amd.fdtd import workloads
from lightsmith_amd.fdtd.batch import Batch
from tests import cases

pytestmark = pytest.mark.gpu

SMALL = dict(n_inst=6, nz=60, size_lo=10, size_hi=40, d=20.0, pml=6, gap=6, pad=3, layer=8, n_angles=3, n_wl=2)


def _sweep(gpu, runs, stride, steps):
    return [workloads.oblique_engine(r, gpu, max_steps=steps + 4 * stride, stride=stride) for r in runs]


def _mixed(gpu, steps):
    """Oblique-sweep members (real TFSF at normal incidence and complex Bloch runs) plus dispersive CPML boxes
    (D-form poles, dipole; indexed and dense materials; 3-pole media, and 2- and 1-pole media in one box), all on one
    DFT stride."""
    runs = workloads.oblique_sweep_scene(**SMALL)
    stride = 10
    engs = _sweep(gpu, runs, stride, steps)
    ns = cases.product_ns()
    engs.append(cases.dispersive_box(ns, gpu, shape=(40, 36, 32), n_steps_hint=steps, material="indexed"))
    engs.append(cases.dispersive_box(ns, gpu, shape=(29, 45, 21), seed=3, n_steps_hint=steps, material="dense"))
    engs.append(cases.dispersive_box(ns, gpu, shape=(24, 30, 20), seed=5, n_steps_hint=steps, material="indexed",
                                     pole_counts=(2, 1)))  # another pole count: its own E launch
    for e in engs:
        assert e.stride == stride
    return engs


def _assert_same(alone, batched):
    for q, (a, b) in enumerate(zip(alone, batched)):
        ta, tb = a.state_tensors(), b.state_tensors()
        assert len(ta) == len(tb)
        for x, y in zip(ta, tb):
            assert x.shape == y.shape and x.dtype == y.dtype
            assert torch.equal(x, y), f"member {q}: a state tensor differs"


def test_batch_members_match_their_runs_alone(gpu):
    chunks, steps = 6, 60
    alone = _mixed(gpu, steps)
    for e in alone:
        for _ in range(chunks):
            e.chunk()
    members = _mixed(gpu, steps)
    b = Batch(members)
    assert b.tiles() > 0
    for _ in range(chunks):
        b.chunk()
    torch.cuda.synchronize()
    _assert_same(alone, members)
    assert all(int(e._counter[0]) == chunks * e.stride for e in members)
    assert any(e.complex for e in members) and any(not e.complex for e in members)
    b.close()


@pytest.mark.parametrize("schedule", [(1, 4, 32, 8, 0), (2, 8, 32, 16, 0), (1, 8, 64, 5, 64), (2, 4, 64, 32, 16)])
def test_batch_schedules_and_graphs_agree(gpu, schedule):
    chunks, steps = 4, 40
    ref = _mixed(gpu, steps)
    rb = Batch(ref)
    for _ in range(chunks):
        rb.chunk()
    members = _mixed(gpu, steps)
    b = Batch(members, schedule=schedule)
    run, used = b.make_runner(True)
    assert used
    for _ in range(chunks):
        run()
    b.release()
    torch.cuda.synchronize()
    _assert_same(ref, members)
    rb.close()
    b.close()


def test_autotune_leaves_no_trace(gpu):
    chunks, steps = 3, 30
    ref = _mixed(gpu, steps)
    rb = Batch(ref)
    members = _mixed(gpu, steps)
    b = Batch(members)
    b.chunk()
    rb.chunk()
    res = b.autotune(candidates=[(1, 4, 64, 16, 0), (1, 4, 32, 8, 0), (2, 8, 64, 16, 2048)], steps=3, warmup=1)
    assert len(res) == 3 and b.schedule == res[0][1]
    for _ in range(chunks):
        b.chunk()
        rb.chunk()
    torch.cuda.synchronize()
    _assert_same(ref, members)
    b.close()
    rb.close()


def test_batched_member_matches_the_oracle(gpu):
    """One dispersive box stepped inside a batch, against the FP64 oracle (the tolerance of the parity
    tests)."""
    steps = 200
    o = cases.dispersive_box(cases.oracle_ns(), shape=(40, 36, 32), n_steps_hint=steps)
    members = _mixed(gpu, steps)
    b = Batch(members)
    for _ in range(steps // o.stride):
        o.chunk()
        b.chunk()
    torch.cuda.synchronize()
    g = members[-3]  # the (40, 36, 32) indexed box
    Eo, Ho = cases.fields(o)
    Eg, Hg = cases.fields(g)
    err = max(cases.rel_l2(Eg, Eo), cases.rel_l2(Hg, Ho))
    eo, _ = o.monitors[0].results()
    eg, _ = g.monitors[0].results()
    derr = cases.rel_l2(eg, eo)
    print(f"batched member vs FP64 oracle: fields {err:.2e}, DFT {derr:.2e}")
    assert err < 1e-5 and derr < 1e-5
    b.close()


def test_batch_refuses_what_it_cannot_run(gpu):
    engs = _mixed(gpu, 20)
    with pytest.raises(ValueError):
        Batch([engs[0], engs[0]])  # one engine twice
    b = Batch(engs[:2])
    with pytest.raises(RuntimeError):
        engs[0].close()  # held by the batch
    b.close()
    odd = workloads.oblique_engine(workloads.oblique_sweep_scene(**SMALL)[1], gpu, max_steps=100, stride=7)
    with pytest.raises(ValueError):
        Batch([engs[2], odd])  # different DFT strides
    np.testing.assert_equal(len(engs), 9)


def test_batch_run_control_matches_sequential_runs(gpu):
    """run_batch_until_decayed stops each member where run_until_decayed stops it alone (staggered caps and
    minimum lengths), with bit-identical monitors."""
    from lightsmith_amd.fdtd.run import RunOptions, run_batch_until_decayed, run_until_decayed
    opts = RunOptions(decay_by=1e-2)
    n = 4
    max_steps = [200 + 40 * q for q in range(n)]
    min_steps = [60 + 20 * q for q in range(n)]
    windows = [3] * n
    runs = workloads.oblique_sweep_scene(**dict(SMALL, n_inst=n))
    alone = _sweep(gpu, runs, 10, 400)
    seq = [run_until_decayed(e, opts, max_steps=max_steps[q], min_steps=min_steps[q], window=windows[q],
                             warnings=[]) for q, e in enumerate(alone)]
    members = _sweep(gpu, runs, 10, 400)
    bat = run_batch_until_decayed(members, opts, max_steps=max_steps, min_steps=min_steps, windows=windows,
                                  warnings=[])
    print("steps alone", [s[0] for s in seq], "batched", [s[0] for s in bat])
    assert [s[0] for s in seq] == [s[0] for s in bat]
    torch.cuda.synchronize()
    _assert_same(alone, members)
