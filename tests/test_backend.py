"""The plugin boundary on the CPU (no solve): ``lightsmith_amd.backend.NativeFDTD`` has the reference backend's class
surface (backends/base.py:20-64, backends/fdtd.py:55-63), declines what the HIP path does not run, plans the
plain Au sphere as the reference lays it out (97^3 nodes, dt = 0.99 * 2 / sqrt(3) nm/c, DFT stride 107), and the
problem types parse the reference's JSON form."""

import json

import numpy as np
import pytest

from lightsmith_amd import api, backend
from lightsmith_amd import problem as P
from lightsmith_amd.fdtd import workloads


def test_class_surface_mirrors_the_reference_backend():
    B = backend.NativeFDTD
    assert (B.name, B.method, B.license, B.gpu) == ("lightsmith-fdtd", "fdtd", "MIT", True)
    assert "scattering" in B.problem_kinds and B.options_model is P.FDTDOptions
    for m in ("available", "supports", "estimate", "solve", "info"):
        assert callable(getattr(B, m))
    assert backend.get_backend("fdtd") is B and backend.get_backend("lightsmith-fdtd") is B
    info = B.info()
    assert info["name"] == "lightsmith-fdtd" and info["available"]


def test_au_sphere_plan_is_the_reference_layout():
    sim = workloads.au_sphere_sim()
    ok, why = backend.NativeFDTD.supports(sim)
    assert ok, why
    est = api.estimate(sim)
    assert est["grid"] == [97, 97, 97] and est["spacing_nm"] == 2.0
    assert est["dft_stride"] == 107
    assert abs(est["dt_fs"] * 299.792458 - 0.99 * 2 / np.sqrt(3)) < 1e-12


@pytest.mark.parametrize("change,fragment", [
    (dict(illumination={"kind": "gaussian_beam", "waist_nm": 500.0}), "illumination"),
    (dict(substrate={"material": {"kind": "constant_index", "n_real": 0.2, "n_imag": 3.0}}), "absorbing substrate"),
    (dict(orientation_average=True), "orientation_average"),
    (dict(options={"fdtd": {"precision": "double"}}), "double"),
    (dict(options={"fdtd": {"grid": "auto"}}), "graded"),
])
def test_unsupported_requests_are_declined_with_a_reason(change, fragment):
    base = json.loads(workloads.au_sphere_sim().model_dump_json())
    base.update(change)
    ok, why = backend.NativeFDTD.supports(base)
    assert not ok and fragment in why


def test_options_are_validated_like_the_reference():
    with pytest.raises(ValueError):
        P.FDTDOptions.model_validate({"spacing_nm": 2.0, "no_such_option": 1})
    with pytest.raises(ValueError):
        P.FDTDOptions.model_validate({"courant": 1.5})


def test_solve_refuses_a_cpu_device():
    with pytest.raises(ValueError, match="GPU"):
        backend.NativeFDTD.solve(workloads.au_sphere_sim(), "cpu")


def test_reference_json_round_trip_and_result_fields():
    sim = workloads.au_sphere_sim(num_wavelengths=5)
    again = P.parse_problem(sim.model_dump_json())
    assert again == sim and np.allclose(again.wavelength_array(), [400, 500, 600, 700, 800])
    r = P.Result.from_arrays(method="fdtd", backend="lightsmith-fdtd", wavelengths_nm=[500.0], c_ext=[3.0],
                             c_sca=[1.0], c_abs=[2.0], r_eq=40.0)
    assert set(["c_ext_nm2", "c_sca_nm2", "c_abs_nm2", "equivalent_radius_nm", "provenance", "warnings"]) <= \
        set(r.model_dump())
    assert np.isclose(r.q_ext[0], 3.0 / (np.pi * 1600))


def test_materials_map_to_the_reference_poles():
    """DrudeLorentz -> Drude pole sigma = (wp / hbar c)^2, w0 = 0, and Lorentz poles sigma = d w0^2 (media.py:68-75);
    the medium's eps(omega) equals the material's eps(lambda)."""
    from lightsmith_amd.fdtd.scattering import to_medium
    mat = P.DrudeLorentz(**workloads.AU_DRUDE_2LORENTZ)
    m = to_medium(mat, np.array([400.0, 800.0]), [])
    assert len(m.poles) == 3 and m.poles[0].omega0 == 0.0
    wl = np.linspace(400, 800, 9)
    assert np.allclose(m.eps(2 * np.pi / wl), mat.eps(wl), rtol=1e-9)


def test_far_field_is_taken_and_declined_over_a_substrate():
    from lightsmith_amd.fdtd import scattering
    base = json.loads(workloads.au_sphere_sim().model_dump_json())
    base["outputs"] = {"far_field": {"theta_deg": [0.0, 90.0], "wavelengths_nm": [550.0, 900.0]}}
    ok, why = backend.NativeFDTD.supports(base)
    assert ok, why
    sim = P.parse_problem(base)
    assert scattering.band_wavelengths(sim).max() == 900.0  # the pulse carries the far field's band too
    base["substrate"] = {"material": {"kind": "constant_index", "n_real": 1.5}, "interface_z_nm": -60.0}
    ok, why = backend.NativeFDTD.supports(base)
    assert not ok and "far_field" in why
