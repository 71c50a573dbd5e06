"""Dipole emission problems on the CPU (no solve): the schema parses the reference's JSON form, the backend takes
emission problems (radiation patterns included, their wavelengths after the results') and declines double precision
with a reason, and the plan follows the reference's rules
(emission.py:46-103): the layout around the particles and the dipole, a quarter of the dipole's gap as the largest
cell, the 6 omega_half stride margin."""

import json

import numpy as np
import pytest

from lightsmith_amd import api, backend
from lightsmith_amd import problem as P
from lightsmith_amd.fdtd import emission


def near_sphere(r=60.0, d=70.0, material=None, n_host=1.0, wl=(450.0, 600.0, 800.0), **fdtd):
    return P.EmissionProblem(particles=[P.Particle(shape=P.Sphere(radius_nm=r),
                                                   material=material or P.ConstantIndex(n_real=2.0))],
                             medium_index=n_host, wavelengths=list(wl),
                             dipole=P.Dipole(position_nm=(0.0, 0.0, d), orientation=(1.0, 0.0, 1.0)),
                             options={"fdtd": fdtd})


def test_emission_problem_parses_and_dispatches():
    p = P.parse_problem({"kind": "emission", "particles": [{"shape": {"kind": "sphere", "radius_nm": 40},
                                                            "material": {"kind": "constant_index", "n_real": 2}}],
                         "dipole": {"position_nm": [0, 0, 50], "orientation": [1, 0, 0]}, "wavelengths": [600]})
    assert isinstance(p, P.EmissionProblem) and np.allclose(p.dipole.unit(), [1, 0, 0])
    assert P.parse_problem(json.loads(p.model_dump_json())) == p
    assert "emission" in backend.NativeFDTD.problem_kinds
    with pytest.raises(ValueError):
        P.Dipole(orientation=(0.0, 0.0, 0.0))


def test_emission_supports_patterns_and_declines_foreign_materials():
    B = backend.NativeFDTD
    assert B.supports(near_sphere())[0]
    assert B.supports(P.EmissionProblem(wavelengths=[500.0]))[0]  # a lone dipole
    pat = near_sphere().model_copy(update={"radiation_pattern": P.RadiationPatternRequest(theta_deg=[0.0],
                                                                                          wavelengths_nm=[600.0])})
    assert B.supports(pat)[0]
    assert emission._wavelengths(pat).tolist() == [450.0, 600.0, 800.0, 600.0]
    bad = near_sphere().model_copy(update={"options": {"fdtd": {"precision": "double"}}})
    ok, why = B.supports(bad)
    assert not ok and "double" in why


def test_emission_plan_follows_the_reference_rules():
    est = api.estimate(near_sphere())
    assert est["runs"] == 2 and est["spacing_nm"] == pytest.approx(2.5)  # 10 nm gap / 4
    p = emission.plan(near_sphere(), P.FDTDOptions())
    assert p["dx"] == pytest.approx(2.5)
    lay = p["layout"]
    z_hi = lay.origin[2] + lay.tfsf[1][2] * p["dx"]
    assert z_hi > 70.0 + 2 * 2.5  # the box reaches past the dipole
    om = 2 * np.pi / np.array([450.0, 600.0, 800.0])
    omega_c = 0.5 * (om.min() + om.max())
    omega_half = max(0.5 * (om.max() - om.min()), 0.15 * omega_c)
    assert p["stride"] == max(1, int(2 * np.pi / (omega_c + 6 * omega_half + om.max()) / p["dt"]))
    lone = emission.plan(P.EmissionProblem(wavelengths=[500.0]), P.FDTDOptions())
    assert lone["dx"] == pytest.approx(500.0 / 20)  # no particle, no gap: the wavelength rule alone
