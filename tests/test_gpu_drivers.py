"""The problem drivers end to end on the GPU (lightsmith_amd/fdtd/drivers.py) against closed forms: a dielectric
sphere's cross sections against Mie theory (oracle/mie.py), an empty periodic cell and a dielectric half-space at
normal and oblique incidence against Fresnel, a dipole's emitted power against the power through a box around
it. The tolerances are discretisation tolerances (a staircase sphere of 12 cells radius; k dx up to 0.16), stated
per test; the measured errors are printed."""

import math

import numpy as np
import pytest

from lightsmith_amd.fdtd import drivers
from lightsmith_amd.fdtd.media import Medium
from tests.test_drivers import thin_film

pytestmark = pytest.mark.gpu


def test_dielectric_sphere_cross_sections_match_mie(gpu):
    from oracle.mie import cross_sections
    r, dx, eps = 60.0, 5.0, 4.0
    wl = np.array([400.0, 500.0, 600.0])
    lay = drivers.scattering_layout((-r, -r, -r), (r, r, r), dx, float(wl.max()))
    host, glass = Medium(1.0), Medium(eps)
    inv, groups = drivers.staircase(lay.shape, lay.origin, dx, host, [drivers.Sphere((0.0, 0.0, 0.0), r, glass)])
    res = drivers.solve_scattering(gpu, lay, inv, groups, [host, glass], wl, khat=(0, 0, 1.0), ehat=(1.0, 0, 0),
                                   opts=drivers.DriverOptions(decay_by=1e-6))
    ce, cs, _ = cross_sections(eps, r, wl)
    e_ext = np.abs(res.c_ext / ce - 1)
    e_sca = np.abs(res.c_sca / cs - 1)
    print(f"sphere r={r} nm, dx={dx}: C_ext {res.c_ext} vs Mie {ce} (rel {e_ext}); C_sca rel {e_sca}; "
          f"C_abs/C_ext {res.c_abs / res.c_ext}; steps {res.provenance['steps']}")
    assert np.all(e_ext < 0.06) and np.all(e_sca < 0.06)
    assert np.all(np.abs(res.c_abs) < 0.01 * res.c_ext)


def _cell(eps_sub, n=(4, 4), d=10.0, pml=12, gap=24, pad=6, layer=10):
    k_r = pml + gap
    k_src = k_r + 2
    k_top = k_src + pad
    k_bot = k_top + layer
    k_t = k_bot + 3
    nz = k_t + gap + pml + 2
    shape = (n[0] + 2, n[1] + 2, nz)
    eps = np.ones(shape)
    eps[:, :, k_bot:] = eps_sub
    inv = np.broadcast_to(1 / eps, (3,) + shape).copy()
    return drivers.PeriodicCell(n_cells=n, nz=nz, spacing=(d, d, d), inv_eps=inv, groups=[], pml=pml, k_src=k_src,
                                k_r=k_r, k_t=k_t, k_top=k_top, eps_sup=1.0, eps_sub=eps_sub)


def test_empty_cell_transmits_everything(gpu):
    res = drivers.solve_periodic(gpu, _cell(1.0), [500.0, 600.0, 700.0])
    R, T = res.r[:, 0, 0], res.t[:, 0, 0]
    print(f"empty cell: R {R}, T {T}")
    assert np.all(np.abs(R) < 1e-3) and np.all(np.abs(T - 1) < 1e-3)
    assert np.all(np.abs(res.r).sum(axis=(1, 2)) < 2e-3)


@pytest.mark.parametrize("pol", ["s", "p"])
def test_half_space_normal_incidence_is_fresnel(gpu, pol):
    n2 = 1.5
    res = drivers.solve_periodic(gpu, _cell(n2 ** 2), [500.0, 600.0, 700.0], pol=pol)
    R, T = res.r[:, 0, 0], res.t[:, 0, 0]
    Rf = ((1 - n2) / (1 + n2)) ** 2
    print(f"half-space {pol}: R {R} T {T} (Fresnel {Rf:.5f}, {1 - Rf:.5f})")
    assert np.all(np.abs(R - Rf) < 4e-3) and np.all(np.abs(T - (1 - Rf)) < 4e-3)


@pytest.mark.parametrize("pol", ["s", "p"])
def test_half_space_oblique_incidence_is_fresnel(gpu, pol):
    n2, theta, lam = 1.5, math.radians(30.0), 600.0
    k = 2 * math.pi / lam * math.sin(theta)
    res = drivers.solve_periodic(gpu, _cell(n2 ** 2), [lam], pol=pol, kpar=(k, 0.0))
    R, T = float(res.r[0, 0, 0]), float(res.t[0, 0, 0])
    ci = math.cos(theta)
    ct = math.sqrt(1 - (math.sin(theta) / n2) ** 2)
    if pol == "s":
        Rf = ((ci - n2 * ct) / (ci + n2 * ct)) ** 2
    else:
        Rf = ((n2 * ci - ct) / (n2 * ci + ct)) ** 2
    print(f"half-space oblique {pol}: R {R:.5f} T {T:.5f} (Fresnel {Rf:.5f})")
    assert abs(R - Rf) < 8e-3 and abs(T - (1 - Rf)) < 8e-3  # the reference's tolerance (test_fdtd_periodic.py:91-103)


def test_dipole_power_leaves_through_the_box(gpu):
    dx = 10.0  # k dx 0.09-0.13: the dipole's quasi-static near field, (k dx)^-3 times the radiated one, stays
    # small enough for float32 fields
    wl = np.array([500.0, 600.0, 700.0])
    lay = drivers.scattering_layout((-30, -30, -30), (30, 30, 30), dx, float(wl.max()))
    host = Medium(1.0)
    inv = np.ones((3,) + lay.shape)
    res = drivers.solve_emission(gpu, lay, inv, [], [host], wl, pos=(1.5, -1.0, 0.5), pdir=(0.0, 0.0, 1.0))
    ratio = res.p_rad / res.p_tot
    print(f"dipole: p_rad / p_tot {ratio}, steps {res.provenance['steps']}")
    assert np.all(res.p_tot > 0) and np.all(np.abs(ratio - 1) < 0.01)


def _film_cell(d, nf, film, eps_sub=2.25, n=(4, 4), pml_nm=80.0, gap_nm=240.0, pad_nm=60.0, above_nm=30.0):
    """A periodic cell with a film of nf cells (thickness nf d) on a substrate: the film's E nodes take 1/eps_inf of
    `film`, or are D-form nodes of it when it has poles; interfaces sit half a cell below the film's first and last E_x/E_y planes, the E_z
    nodes (half a cell up) take the medium at their own position."""
    from lightsmith_amd.fdtd.engine import DispersiveGroup
    pml, gap, pad = (int(round(v / d)) for v in (pml_nm, gap_nm, pad_nm))
    layer = nf + int(round(above_nm / d))  # the incident plane k_top sits above_nm above the film
    k_r = pml + gap
    k_src = k_r + 2
    k_top = k_src + pad
    k_bot = k_top + layer
    k_t = k_bot + 3
    nz = k_t + gap + pml + 2
    shape = (n[0] + 2, n[1] + 2, nz)
    inv = np.ones((3,) + shape)
    film_sel = np.zeros((3,) + shape, bool)
    inv[0:2, :, :, k_bot:] = 1 / eps_sub
    inv[2, :, :, k_bot - 1:] = 1 / eps_sub
    film_sel[0:2, :, :, k_bot - nf:k_bot] = True
    film_sel[2, :, :, k_bot - nf - 1:k_bot - 1] = True
    groups = []
    if film.poles:
        inv[film_sel] = 1.0
        idx = np.nonzero(film_sel.reshape(-1))[0].astype(np.int64)
        groups.append(DispersiveGroup(idx=idx, media=[film], frac=np.ones((1, len(idx))), w=None))
    else:
        inv[film_sel] = 1.0 / film.eps_inf
    return drivers.PeriodicCell(n_cells=n, nz=nz, spacing=(d, d, d), inv_eps=inv, groups=groups, pml=pml, k_src=k_src,
                                k_r=k_r, k_t=k_t, k_top=k_top, eps_sup=1.0, eps_sub=eps_sub)


def test_drude_film_is_the_thin_film_formula(gpu):
    """A 30 nm Drude film (eps_inf 5, wp 9 eV, gamma 0.05 eV) on glass, as the reference's layer-stack tests set it
    (test_fdtd_periodic.py:55-63 at normal incidence, tolerance 0.005; :120-126 at 30 and 60 deg, 0.003), against the
    three-layer closed form with the continuous Drude permittivity, 2.5 nm cells. Normal incidence: one broadband real
    run; oblique: one complex Bloch run per angle and polarisation. Measured: 4e-4 at normal incidence, 2e-3 at 60 deg p
    (1.4e-3 and 3.3e-3 with 5 nm cells)."""
    from lightsmith_amd.fdtd.media import drude_lorentz
    film = drude_lorentz(5.0, 9.0, 0.05)
    d, nf = 2.5, 12
    cell = _film_cell(d, nf, film)
    n_sub = 1.5
    wl = np.array([450.0, 600.0, 750.0])
    res = drivers.solve_periodic(gpu, cell, wl, pol="s")
    worst = 0.0
    for q, lam in enumerate(wl):
        Rt, Tt = thin_film(1.0, np.sqrt(film.eps(2 * np.pi / lam)), n_sub, nf * d, lam, 0.0, "s")
        R, T = float(res.r[q, 0, 0]), float(res.t[q, 0, 0])
        print(f"Drude film, normal, {lam:.0f} nm: R {R:.5f} (TMM {Rt:.5f}), T {T:.5f} (TMM {Tt:.5f})")
        worst = max(worst, abs(R - Rt), abs(T - Tt))
    assert worst < 0.005
    lam = 600.0
    worst = 0.0
    for deg in (30.0, 60.0):
        th = math.radians(deg)
        for pol in "sp":
            r = drivers.solve_periodic(gpu, cell, [lam], pol=pol, kpar=(2 * math.pi / lam * math.sin(th), 0.0))
            Rt, Tt = thin_film(1.0, np.sqrt(film.eps(2 * np.pi / lam)), n_sub, nf * d, lam, th, pol)
            R, T = float(r.r[0, 0, 0]), float(r.t[0, 0, 0])
            print(f"Drude film, {deg:.0f} deg {pol}: R {R:.5f} (TMM {Rt:.5f}), T {T:.5f} (TMM {Tt:.5f})")
            worst = max(worst, abs(R - Rt), abs(T - Tt))
    assert worst < 0.003


def test_dielectric_slab_is_the_thin_film_formula(gpu):
    """A 300 nm n = 2 slab on glass, the reference's layer-stack case (test_fdtd_periodic.py:55-63: 10 nm cells,
    400-800 nm, tolerance 0.01 at normal incidence; :91-103: 450 and 750 nm at 30 and 60 deg, s and p, 0.008),
    against the three-layer closed form."""
    d, nf = 10.0, 30
    cell = _film_cell(d, nf, Medium(4.0))
    wl = np.linspace(400.0, 800.0, 9)
    res = drivers.solve_periodic(gpu, cell, wl, pol="s")
    worst = 0.0
    for q, lam in enumerate(wl):
        Rt, Tt = thin_film(1.0, 2.0, 1.5, nf * d, lam, 0.0, "s")
        R, T = float(res.r[q, 0, 0]), float(res.t[q, 0, 0])
        worst = max(worst, abs(R - Rt), abs(T - Tt))
    print(f"n = 2 slab, normal incidence: max |dR|, |dT| {worst:.4f}")
    assert worst < 0.01
    worst = 0.0
    for deg in (30.0, 60.0):
        th = math.radians(deg)
        for lam in (450.0, 750.0):
            for pol in "sp":
                r = drivers.solve_periodic(gpu, cell, [lam], pol=pol, kpar=(2 * math.pi / lam * math.sin(th), 0.0))
                Rt, Tt = thin_film(1.0, 2.0, 1.5, nf * d, lam, th, pol)
                worst = max(worst, abs(float(r.r[0, 0, 0]) - Rt), abs(float(r.t[0, 0, 0]) - Tt))
                print(f"n = 2 slab, {deg:.0f} deg {lam:.0f} nm {pol}: R {float(r.r[0, 0, 0]):.5f} (TMM {Rt:.5f})")
    assert worst < 0.008
