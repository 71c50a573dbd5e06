"""Host logic of the problem drivers (lightsmith_amd/fdtd/drivers.py): layout, band, voxelisation, periodic time
step. The runs themselves are GPU tests (test_gpu_drivers.py)."""

import math

import numpy as np

from lightsmith_amd.fdtd import drivers
from lightsmith_amd.fdtd.media import Medium, Pole


def test_scattering_layout_follows_the_reference_rule():
    lay = drivers.scattering_layout((-60, -60, -60), (60, 60, 60), 5.0, 600.0)
    # half_cells = ceil(60 / 5) + 1 = 13; padding and gap max(4, ceil(0.03 * 600 / 5)) = 4; 10 CPML cells
    assert lay.shape == (69, 69, 69) and lay.pml == 10
    c = lay.center
    assert np.all(c == 34)
    assert np.all(lay.tfsf[0] == c - 17) and np.all(lay.tfsf[1] == c + 17)
    assert np.all(lay.tot[0] == lay.tfsf[0] + 2) and np.all(lay.sca[1] == lay.tfsf[1] + 2)
    assert np.all(lay.sca[0] - 4 - 1 == lay.pml)  # gap, then one node, then the CPML
    assert np.allclose(lay.origin + c * lay.dx, 0.0)
    wide = drivers.scattering_layout((0, 0, 0), (10, 10, 10), 2.0, 2000.0, n_host=1.5)
    # the 3 % floor: 0.03 * 2000 / 1.5 = 40 nm = 20 cells of padding and gap
    assert np.all(wide.tfsf[0] - wide.sca[0] == 2) and np.all(wide.tfsf[1] - wide.center == 4 + 20)


def test_band_covers_the_wavelengths():
    oc, oh, om = drivers.band([400.0, 600.0])
    w = 2 * np.pi / np.array([400.0, 600.0])
    assert math.isclose(oc, w.mean()) and math.isclose(oh, 0.5 * (w[0] - w[1])) and math.isclose(om, w[0])
    oc, oh, _ = drivers.band([500.0])
    assert math.isclose(oh, 0.15 * oc)


def test_staircase_sphere_volume_and_material_kinds():
    shape, dx = (41, 41, 41), 2.0
    origin = -20 * dx * np.ones(3)
    host = Medium(1.0)
    glass = Medium(2.25)
    inv, groups = drivers.staircase(shape, origin, dx, host, [drivers.Sphere((0, 0, 0), 20.0, glass)])
    n_in = np.sum(inv == 1 / 2.25)
    vol = 4 / 3 * math.pi * (20.0 / dx) ** 3
    assert abs(n_in / (3 * vol) - 1) < 0.02 and not groups
    drude = Medium(1.0, (Pole(0.01, 0.0, 0.001),))
    inv2, groups2 = drivers.staircase(shape, origin, dx, host, [drivers.Sphere((0, 0, 0), 20.0, drude)])
    (g,) = groups2
    assert len(g.idx) == n_in and np.all(inv2.reshape(-1)[g.idx] == 1.0)
    # the group's nodes are exactly the E nodes inside the sphere
    c, i, j, k = np.unravel_index(g.idx, (3,) + shape)
    pos = origin[None, :] + (np.stack([i, j, k], 1) + 0.5 * np.eye(3)[c]) * dx
    assert np.all(np.sum(pos ** 2, axis=1) <= 400.0 + 1e-9)


def test_periodic_time_step_single_cell_axis_uses_the_bloch_eigenvalue():
    inv = np.ones((3, 3, 3, 40))
    cell = drivers.PeriodicCell(n_cells=(1, 1), nz=40, spacing=(10.0, 10.0, 10.0), inv_eps=inv, groups=[], pml=8,
                                k_src=12, k_r=10, k_t=30, k_top=14)
    media = [Medium(1.0)]
    assert math.isclose(drivers.periodic_time_step(cell, 0.9, media), 0.9 * 10.0)
    k = 0.02
    dt = drivers.periodic_time_step(cell, 0.9, media, (k, 0.0))
    assert math.isclose(dt, 0.9 / math.sqrt(1 / 100 + (math.sin(k * 5.0) / 10.0) ** 2))


def thin_film(n1, n2, n3, t, lam, theta, pol):
    """(R, T) of a film (index n2, thickness t) between two half-spaces, s or p, incidence angle theta in medium 1:
    the closed-form three-layer Airy sum (e^{-i omega t} convention, Im n2 >= 0)."""
    s1 = n1 * math.sin(theta)
    n = (n1, n2, n3)
    c = [np.sqrt(1 - (s1 / m) ** 2 + 0j) for m in n]

    def r(a, b):
        if pol == "s":
            return (n[a] * c[a] - n[b] * c[b]) / (n[a] * c[a] + n[b] * c[b])
        return (n[b] * c[a] - n[a] * c[b]) / (n[b] * c[a] + n[a] * c[b])

    def tr(a, b):
        if pol == "s":
            return 2 * n[a] * c[a] / (n[a] * c[a] + n[b] * c[b])
        return 2 * n[a] * c[a] / (n[b] * c[a] + n[a] * c[b])

    e2 = np.exp(4j * np.pi * n2 * c[1] * t / lam)  # round trip through the film
    den = 1 + r(0, 1) * r(1, 2) * e2
    rr = (r(0, 1) + r(1, 2) * e2) / den
    tt = tr(0, 1) * tr(1, 2) * np.sqrt(e2) / den
    fac = (n3 * c[2]).real / (n1 * c[0]).real if pol == "s" else (np.conj(n3) * c[2]).real / (np.conj(n1) * c[0]).real
    return float(abs(rr) ** 2), float(fac * abs(tt) ** 2)


def test_tmm_closed_form_conserves_energy_and_reduces_to_fresnel():
    """The test's own closed form: lossless films conserve energy; a film of the substrate's index is a bare
    interface (Fresnel); a quarter-wave layer of index sqrt(n1 n3) reflects nothing, a half-wave layer is absent
    (these two pin the film's phase)."""
    for pol in "sp":
        for th in (0.0, 0.5, 1.0):
            R, T = thin_film(1.0, 2.0, 1.5, 30.0, 600.0, th, pol)
            assert abs(R + T - 1) < 1e-12
            R, _ = thin_film(1.0, 1.5, 1.5, 30.0, 600.0, th, pol)
            ci, ct = math.cos(th), math.sqrt(1 - (math.sin(th) / 1.5) ** 2)
            Rf = ((ci - 1.5 * ct) / (ci + 1.5 * ct)) ** 2 if pol == "s" else ((1.5 * ci - ct) / (1.5 * ci + ct)) ** 2
            assert abs(R - Rf) < 1e-12
    n2 = math.sqrt(1.5)
    R, T = thin_film(1.0, n2, 1.5, 600.0 / (4 * n2), 600.0, 0.0, "s")
    assert R < 1e-20 and abs(T - 1) < 1e-12
    R, _ = thin_film(1.0, 2.0, 1.5, 600.0 / (2 * 2.0), 600.0, 0.0, "p")
    assert abs(R - 0.04) < 1e-12  # ((1 - 1.5) / (1 + 1.5))^2

