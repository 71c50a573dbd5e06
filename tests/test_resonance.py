"""The product's harmonic inversion (lightsmith_amd/fdtd/resonance.py, the reference's resonance.py) on synthetic
ring-downs whose modes are known exactly."""

import numpy as np


def test_harmonic_inversion_recovers_synthetic_modes():
    """lightsmith_amd.fdtd.resonance: three decaying modes (Q 8, 35, 120) in 6 channels with different amplitudes
    and phases, sampled well above the band with a few 1e-6 of noise: frequencies to 1e-6, Q to 1e-4, the
    amplitude order; an out-of-band mode and a growing one are not reported."""
    from lightsmith_amd.fdtd.resonance import harmonic_inversion
    rng = np.random.default_rng(11)
    dt = 0.05
    t = np.arange(6000) * dt
    modes = [(1.2, 8.0, 1.0), (1.5, 35.0, 0.5), (1.9, 120.0, 0.2), (3.0, 50.0, 0.7)]  # the last out of band
    x = np.zeros((6, len(t)))
    for w, q, a in modes:
        ph = rng.uniform(0, 2 * np.pi, 6)
        amp = a * rng.uniform(0.5, 1.5, 6)
        x += (amp[:, None] * np.exp(-w / (2 * q) * t)[None, :] * np.cos(w * t[None, :] + ph[:, None]))
    x += 3e-6 * rng.standard_normal(x.shape)
    found = harmonic_inversion(x, dt, 1.0, 2.2, max_modes=10)
    got = sorted((m.omega.real, m.q) for m in found)
    assert len(got) == 3
    for (w, q, _), (gw, gq) in zip(modes[:3], got):
        assert abs(gw / w - 1) < 1e-6 and abs(gq / q - 1) < 1e-4
    assert [round(m.omega.real, 3) for m in found] == [1.2, 1.5, 1.9]  # strongest first
