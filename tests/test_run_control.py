"""Run control (setup.py:722-773) and flux geometry (setup.py:432-468, 564-566) on the CPU."""

import math

import numpy as np
import pytest

from lightsmith_amd.fdtd import run as R


class FakeEngine:
    """An engine whose monitor energy follows a given sequence, one value per chunk."""

    def __init__(self, energies, stride=10):
        self.seq = list(energies)
        self.stride = stride
        self.energy = np.zeros(2)
        self.released = False
        self.k = 0

    def make_runner(self, use_graphs):
        def run():
            self.energy = np.array([self.seq[min(self.k, len(self.seq) - 1)], 0.0])
            self.k += 1
        return run, use_graphs

    def release(self):
        self.released = True


def test_decay_shutoff_after_min_steps_and_window():
    pulse = [math.exp(-((k - 10) / 3.0) ** 2) for k in range(200)]
    eng = FakeEngine(pulse)
    steps, level, graphs = R.run_until_decayed(eng, R.RunOptions(decay_by=1e-5), max_steps=2000, min_steps=150,
                                               window=4, warnings=[])
    assert eng.released and graphs
    assert 150 <= steps < 400 and level < 1e-5


def test_divergence_raises_and_releases():
    for seq in ([1.0] * 5 + [float("nan")], [1.0] * 3 + [10.0 ** k for k in range(30)]):
        eng = FakeEngine(seq)
        with pytest.raises(FloatingPointError, match="diverged"):
            R.run_until_decayed(eng, R.RunOptions(), max_steps=10_000, min_steps=20, window=3, warnings=[])
        assert eng.released


def test_time_cap_warns_and_fixed_run_time_does_not():
    warn = []
    R.run_until_decayed(FakeEngine([1.0] * 100), R.RunOptions(), max_steps=200, min_steps=10, window=3, warnings=warn)
    assert warn and "decayed only" in warn[0]
    warn = []
    steps, _, _ = R.run_until_decayed(FakeEngine([1.0] * 100), R.RunOptions(run_time_fs=1.0), max_steps=200,
                                      min_steps=10, window=3, warnings=warn)
    assert steps == 200 and not warn


def test_progress_events_and_cancellation():
    events = []
    with R.reporting(events.append):
        R.run_until_decayed(FakeEngine([1.0] * 10), R.RunOptions(run_time_fs=1.0), max_steps=50, min_steps=10,
                            window=3, warnings=[])
    assert [e["steps"] for e in events] == [10, 20, 30, 40, 50] and all(e["stage"] == "fdtd_steps" for e in events)

    def cancel(ev):
        if ev["steps"] >= 20:
            raise R.Cancelled()

    eng = FakeEngine([1.0] * 10)
    with R.reporting(cancel), pytest.raises(R.Cancelled):
        R.run_until_decayed(eng, R.RunOptions(), max_steps=500, min_steps=10, window=3, warnings=[])
    assert eng.released


def test_flux_box_of_a_dipole_carries_its_power():
    """Poynting's theorem on the grid: in a lossless medium the power a point dipole does on the field,
    -1/2 dx^3 Re(E(x_src) . J*), leaves through a closed flux box around it. Pins the box geometry (faces,
    staggered E/H pairing, trapezoid weights, signs) together with the DFT monitors and the dipole spectrum
    (emission.py:186-190)."""
    from oracle import fdtd, sources

    shape, dx = (36, 36, 36), 1.0
    dt = 0.95 / math.sqrt(3)
    inv = np.ones((3,) + shape)
    omega = np.array([0.35, 0.45, 0.55])
    src = sources.PointDipole(shape=shape, origin=np.zeros(3), dx=dx, dt=dt, pos=(18.0, 17.6, 18.3),
                              pdir=(0.3, 0.2, 0.9), inv_eps=inv, omega_c=0.45, omega_half=0.15, max_steps=2000)
    box = R.flux_points(np.array([10, 10, 10]), np.array([26, 26, 26]), shape, dx)
    m_src = fdtd.DFTMonitor(omega, src.idx)
    m_box = fdtd.DFTMonitor(omega, box[0], box[1], box[2], energy=1)
    stride = 4
    eng = fdtd.Engine(shape=shape, dx=dx, dt=dt, inv_eps=inv, pml=fdtd.PMLSpec(cells=((6, 6),) * 3),
                      sources=[src], monitors=[m_src, m_box], dft_stride=stride)
    n_chunks = 300
    for _ in range(n_chunks):
        eng.chunk()
    j = np.exp(1j * np.outer(omega, (stride * np.arange(1, n_chunks + 1) - 0.5) * dt)) @ src.wave[
        stride * np.arange(1, n_chunks + 1) - 1]
    e_src, _ = m_src.results()
    p_tot = -0.5 * dx ** 3 * np.real((e_src @ src.w) * np.conj(j))
    eb, hb = m_box.results()
    p_rad = R.flux(eb, hb, box[3])
    assert np.all(p_tot > 0)
    assert np.abs(p_rad / p_tot - 1).max() < 5e-3
    # a flipped face sign or an unaveraged H breaks the balance
    assert np.abs(R.flux(eb, hb, -box[3]) / p_tot - 1).max() > 1
