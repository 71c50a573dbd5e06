"""Pins of the oracle's time monitor and of the energy proxy that drives decay termination (engine.py:213-224,
568-592 as restated in oracle/fdtd.py), by the closed-form evolution of a PEC cavity mode.

TM110 of a PEC box, E_z = X(i) Y(j) with X = sin(kx i), kx = pi / (n0 - 1) (Y likewise), started from rest
(H^{-1/2} = 0), evolves under the Yee update as

    E^n = mode cos((n + 1/2) t) / cos(t / 2),     H^{n-1/2} = -cH curl(mode) sin(n t) / sin t,

cos t = 1 - 2 s^2, s^2 = dt^2 / (4 eps) (Kx^2 + Ky^2), K = 2 sin(k d / 2) / d (the leapfrog recurrence
E^{n+1} = 2 cos t E^n - E^{n-1} with E^1 = (2 cos t - 1) E^0, and H summed from it). The forward differences of
the mode are closed forms too, X(i+1) - X(i) = 2 sin(kx / 2) cos(kx (i + 1/2)), so every sample the monitors take
is known exactly. An off-by-one sample position, a wrong averaging partner or a dropped H term fails (planted
below).
"""

import numpy as np
import pytest

from oracle import fdtd

SHAPE, EPS, DX = (17, 13, 6), 2.0, 1.5
DT = 0.9 * DX / np.sqrt(3)
KX, KY = np.pi / (SHAPE[0] - 1), np.pi / (SHAPE[1] - 1)


def _theta():
    K2 = (2 * np.sin(KX / 2) / DX) ** 2 + (2 * np.sin(KY / 2) / DX) ** 2
    return np.arccos(1 - 2 * DT ** 2 / (4 * EPS) * K2)


def _cavity(stride, monitors=(), time_monitors=()):
    eng = fdtd.Engine(shape=SHAPE, dx=DX, dt=DT, inv_eps=np.full((3,) + SHAPE, 1 / EPS), monitors=monitors,
                      time_monitors=time_monitors, dft_stride=stride)
    i, j = np.arange(SHAPE[0]), np.arange(SHAPE[1])
    eng.E[2][:, :, :-1] = np.outer(np.sin(KX * i), np.sin(KY * j))[:, :, None]
    return eng


E_PTS = [(5, 4, 2), (9, 7, 0), (12, 3, 4)]


def _e_expected(n):
    t = _theta()
    return np.array([np.sin(KX * i) * np.sin(KY * j) for i, j, _ in E_PTS]) * np.cos((n + 0.5) * t) / np.cos(t / 2)


def test_time_monitor_records_E_at_the_end_of_each_chunk():
    stride, n_samp = 7, 12
    tm = fdtd.TimeMonitor(np.array([fdtd.flat_index(2, *p, SHAPE) for p in E_PTS]), n_samp)
    eng = _cavity(stride, time_monitors=[tm])
    for _ in range(n_samp):
        eng.chunk()
    got = tm.results(n_samp)  # (channels, samples)
    exp = np.stack([_e_expected((p + 1) * stride) for p in range(n_samp)], axis=1)
    assert np.abs(got - exp).max() < 1e-12


def test_time_monitor_pin_fails_on_a_planted_off_by_one():
    stride, n_samp = 7, 12
    tm = fdtd.TimeMonitor(np.array([fdtd.flat_index(2, *p, SHAPE) for p in E_PTS]), n_samp + 1)
    eng = _cavity(stride, time_monitors=[tm])
    tm.pos = 1  # planted: the first sample lands one slot late
    for _ in range(n_samp):
        eng.chunk()
    exp = np.stack([_e_expected((p + 1) * stride) for p in range(n_samp)], axis=1)
    assert np.abs(tm.results(n_samp) - exp).max() > 1e-2


# H pairs averaged onto E_z nodes: H_x at (i, j) and (i, j - 1); H_y at (i, j) and (i - 1, j)
HX_PTS = [(5, 4, 2), (9, 7, 1)]
HY_PTS = [(12, 3, 4), (6, 8, 3)]


def _monitor(energy=1):
    e = np.array([fdtd.flat_index(2, *p, SHAPE) for p in E_PTS])
    h1 = [fdtd.flat_index(0, i, j, k, SHAPE) for i, j, k in HX_PTS] + \
         [fdtd.flat_index(1, i, j, k, SHAPE) for i, j, k in HY_PTS]
    h2 = [fdtd.flat_index(0, i, j - 1, k, SHAPE) for i, j, k in HX_PTS] + \
         [fdtd.flat_index(1, i - 1, j, k, SHAPE) for i, j, k in HY_PTS]
    # as many E as H points (a flux plane's layout); the E side repeats the first point
    e = np.concatenate([e, e[:1]])
    return fdtd.DFTMonitor(np.array([0.1]), e, np.array(h1), np.array(h2), energy=energy)


def _energy_expected(n):
    t = _theta()
    cH = DT / DX
    se = _e_expected(n)
    se = np.concatenate([se, se[:1]])
    S = np.sin(n * t) / np.sin(t)
    # average of -cH X(i) (Y(j+1) - Y(j)) over j and j - 1: -cH X(i) sin(ky) cos(ky j)
    hx = [-cH * np.sin(KX * i) * np.sin(KY) * np.cos(KY * j) * S for i, j, _ in HX_PTS]
    # average of +cH (X(i+1) - X(i)) Y(j) over i and i - 1: cH sin(kx) cos(kx i) Y(j)
    hy = [cH * np.sin(KX) * np.cos(KX * i) * np.sin(KY * j) * S for i, j, _ in HY_PTS]
    sh = np.array(hx + hy)
    return np.sum(se ** 2) + np.sum(sh ** 2)


@pytest.mark.parametrize("group", [1, 2])
def test_energy_proxy_is_the_sum_of_squared_samples(group):
    stride = 5
    mon = _monitor(energy=group)
    eng = _cavity(stride, monitors=[mon])
    for p in range(15):
        eng.chunk()
        n = (p + 1) * stride
        assert eng.n == n
        exp = _energy_expected(n)
        assert abs(eng.energy[group - 1] - exp) < 1e-12 * max(exp, 1e-3)
        assert eng.energy[2 - group] == 0


def test_energy_pin_fails_on_a_planted_averaging_slip():
    stride = 5
    mon = _monitor()
    mon.h_idx2 = mon.h_idx1  # planted: no average over the two H nodes
    eng = _cavity(stride, monitors=[mon])
    errs = []
    for p in range(15):
        eng.chunk()
        exp = _energy_expected((p + 1) * stride)
        errs.append(abs(eng.energy[0] - exp) / exp)
    assert max(errs) > 1e-3
