"""The table-fed plane wave's tables (lightsmith_amd.fdtd.layered, the construction half of the reference's
ObliqueLayeredPlaneWave) pinned on the CPU: fed to the FP64 oracle's table source in an empty grid, the box must
not leak (the tables are the grid's own discrete wave: the reference's null-particle test, test_fdtd.py:179-200),
and inside the box the field must be the plane wave that ``background`` describes (test_fdtd.py:194-200, < 1e-3).
A planted slip in the discrete wave (the continuum kz instead of the grid's) leaks."""

import numpy as np
import pytest

from lightsmith_amd.fdtd import layered
from lightsmith_amd.fdtd.media import Medium
from oracle import fdtd, sources

SHAPE = (34, 32, 36)
LO, HI = np.array([10, 9, 11]), np.array([23, 22, 25])
DX, EPS = 4.0, 1.69
DT = 0.95 * DX / np.sqrt(3) * np.sqrt(EPS)


def _run(khat, ehat, steps=None, monitors=(), src=None):
    host = Medium(EPS)
    wl = np.array([260.0, 320.0, 400.0])
    om = 2 * np.pi / wl
    omega_c, omega_half = 0.5 * (om.min() + om.max()), max(0.5 * (om.max() - om.min()), 0.15 * 0.5 * (om.min() + om.max()))
    inv = np.full((3,) + SHAPE, 1 / EPS)
    if src is None:
        src = layered.ObliqueLayeredPlaneWave(shape=SHAPE, origin=np.zeros(3), spacing=DX, dt=DT, lo=LO, hi=HI,
                                              khat=khat, ehat=ehat, inv_eps=inv, host=host, omega_c=omega_c,
                                              omega_half=omega_half, max_steps=4000, omega_min=om.min())
    ts = sources.TableSource(src.T.reshape(-1), src.m, src.cmax, src.points, inc=src.inc,
                               T_inc=np.asarray(src.T_inc).reshape(-1))
    eng = fdtd.Engine(shape=SHAPE, dx=DX, dt=DT, inv_eps=inv, pml=fdtd.PMLSpec(cells=((6, 6),) * 3, eps_ref=EPS),
                      sources=[ts], monitors=list(monitors), dft_stride=1)
    n = steps or int(np.ceil(src.t_passed / DT)) + 10
    return eng, src, n


def _inside(a):
    return tuple(slice(int(LO[d]) + 1, int(HI[d])) for d in range(3))


def _leak(khat, ehat, src=None):
    """Largest |E| in the scattered-field region (outside the box, 2 cells off it, clear of the CPML) relative to
    the largest |E| inside, over the whole pulse."""
    eng, src, n = _run(khat, ehat, src=src)
    out_max = in_max = 0.0
    mask = np.ones(SHAPE, bool)
    mask[tuple(slice(int(LO[d]) - 2, int(HI[d]) + 3) for d in range(3))] = False
    mask[:8], mask[-8:], mask[:, :8], mask[:, -8:], mask[:, :, :8], mask[:, :, -8:] = (False,) * 6
    for _ in range(n):
        eng.step()
        E2 = np.sum(eng.E ** 2, axis=0)
        out_max = max(out_max, float(E2[mask].max()))
        in_max = max(in_max, float(E2[_inside(0)].max()))
    return np.sqrt(out_max / in_max), src


@pytest.mark.parametrize("khat,ehat", [
    ((0.0, 0.0, 1.0), (1.0, 0.0, 0.0)),
    ((0.35, -0.4, 0.847), (0.0, 0.904, 0.427)),
    ((-0.9, 0.2, 0.3873), (0.4303, 0.0, 0.9027)),
])
def test_table_fed_box_does_not_leak(khat, ehat):
    khat = np.asarray(khat) / np.linalg.norm(khat)
    ehat = np.asarray(ehat) - (np.asarray(ehat) @ khat) * khat
    leak, _ = _leak(khat, ehat)
    print(f"khat {np.round(khat, 3)}: leak {leak:.2e}")
    assert leak < 1e-7


def test_table_fed_box_holds_the_plane_wave_it_describes():
    khat = np.array([0.35, -0.4, 0.847]) / np.linalg.norm([0.35, -0.4, 0.847])
    ehat = np.array([0.0, 0.904, 0.427])
    ehat = ehat - (ehat @ khat) * khat
    wl = np.array([280.0, 340.0])
    omega = 2 * np.pi / wl
    ii, jj, kk = np.meshgrid(np.arange(13, 20, 3), np.arange(12, 19, 3), np.arange(14, 22, 3), indexing="ij")
    pts = np.concatenate([fdtd.flat_index(c, ii.ravel(), jj.ravel(), kk.ravel(), SHAPE) for c in range(3)])
    mon = fdtd.DFTMonitor(omega, pts)
    ref = fdtd.DFTMonitor(omega, np.arange(3))
    eng, src, n = _run(khat, ehat, monitors=[mon])
    ref.source = eng.sources[0]
    eng.monitors.append(ref)
    ref.acc_e = np.zeros((len(omega), 3), complex)
    for _ in range(n):
        eng.chunk()
    got, _ = mon.results()
    e_ref, _ = ref.results()
    unit = src.incident_reference(omega)
    amp = np.sum(e_ref * np.conj(unit), axis=1) / np.sum(np.abs(unit) ** 2, axis=1)
    exp = amp[:, None] * src.background(omega, pts, "E")
    err = np.abs(got - exp).max() / np.abs(exp).max()
    print(f"near field vs the described plane wave: {err:.2e}")
    assert err < 1e-3


def test_leak_pin_fails_with_the_continuum_wavenumber(monkeypatch):
    khat = np.array([0.35, -0.4, 0.847]) / np.linalg.norm([0.35, -0.4, 0.847])
    ehat = np.array([0.0, 0.904, 0.427])
    ehat = ehat - (ehat @ khat) * khat
    orig = layered.ObliqueLayeredPlaneWave._wave

    def continuum(self, w):
        A, kz = orig(self, w)
        w = np.asarray(w, dtype=float)
        return A, w * self.n_inc * self.khat[2] + 0 * kz  # planted: kz of the continuum, not of the grid

    monkeypatch.setattr(layered.ObliqueLayeredPlaneWave, "_wave", continuum)
    leak, _ = _leak(khat, ehat)
    assert leak > 1e-6
