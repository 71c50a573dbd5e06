"""The far field of the scattering path on the MI355X: dC_sca/dOmega of a dielectric sphere from the scattered-field
box's near-to-far transform against Mie's amplitudes (oracle/mie.py), at the reference's tolerances
(test_fdtd.py:304-321: median 2e-3, max 1e-2 over 37 angles in the E- and H-planes; the phi-averaged pattern
integrating to C_sca within 5e-3); a particle of the host medium has no far field."""

import numpy as np
import pytest

from lightsmith_amd import api
from lightsmith_amd import problem as P
from oracle.mie import dcsca_domega

pytestmark = pytest.mark.gpu


def test_sphere_far_field_matches_mie(gpu):
    wl = [500.0, 700.0]
    theta = list(np.linspace(0, 180, 37))
    sim = P.Simulation(particles=[P.Particle(shape=P.Sphere(radius_nm=60), material=P.ConstantIndex(n_real=1.5))],
                       wavelengths=wl, options={"fdtd": {"spacing_nm": 4.0}},
                       outputs=P.Outputs(far_field=P.FarFieldRequest(theta_deg=theta, phi_deg=[0.0, 90.0],
                                                                     wavelengths_nm=wl)))
    res = api.solve(sim, "fdtd", gpu)
    t = np.radians(theta)
    for i, lam in enumerate(wl):
        a = np.asarray(res.far_field[i].dcsca_domega_nm2)
        b = dcsca_domega(2.25, 60.0, lam, theta, [0.0, 90.0])
        rel = np.abs(a / b - 1)
        c = 2 * np.pi * np.trapezoid(0.5 * (a[:, 0] + a[:, 1]) * np.sin(t), t)
        print(f"{lam} nm: dC/dOmega vs Mie median {np.median(rel):.2e}, max {rel.max():.2e}; integral / C_sca "
              f"{c / res.c_sca_nm2[i]:.5f}")
        assert np.median(rel) < 2e-3 and rel.max() < 1e-2
        assert abs(c / res.c_sca_nm2[i] - 1) < 5e-3


def test_null_particle_has_no_far_field(gpu):
    sim = P.Simulation(particles=[P.Particle(shape=P.Sphere(radius_nm=20), material=P.ConstantIndex(n_real=1.33))],
                       medium_index=1.33, wavelengths=[500.0, 700.0],
                       illumination=P.PlaneWave(direction=(1.0, 2.0, 3.0), polarization=(3.0, 0.0, -1.0)),
                       options={"fdtd": {"spacing_nm": 4.0}},
                       outputs=P.Outputs(far_field=P.FarFieldRequest(theta_deg=[0.0, 90.0], wavelengths_nm=[600.0])))
    res = api.solve(sim, "fdtd", gpu)
    ff = np.abs(np.asarray(res.far_field[0].dcsca_domega_nm2)).max()
    print(f"null particle: max dC/dOmega {ff:.2e} nm^2/sr, C_sca {res.c_sca_nm2}")
    assert ff < 1e-8
