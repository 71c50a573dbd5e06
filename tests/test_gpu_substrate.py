"""Particles on a glass substrate solved on the MI355X (lightsmith_amd/fdtd/scattering.py with a substrate):

* the HIP engine of a metal sphere resting on glass, lit at 40 deg (p) from above -- planar background, vertex
  tensor of the particle, source tables over the substrate, flux boxes -- against the FP64 oracle on the same inputs
  for 3 DFT chunks from mid-pulse: fields and every monitor spectrum within 1e-5, or within 1.25 times the float32
  floor where that is higher (the oracle itself run in float32 on the same case, measured alongside: a metal under
  an abrupt switch-on, as for the Au sphere);
* C_sca of an n = 2, r = 20 nm sphere 40 nm above glass against the dipole model (oracle/dipole_substrate.py) at
  the reference's tolerances (test_fdtd_substrate.py:386-420): from above at normal incidence and at 45 deg s and
  p, through the glass at normal incidence and at 30 deg p;
* the reference's null-particle tests (test_fdtd_substrate.py:327-349): a sphere of the host medium on glass
  scatters nothing and absorbs nothing, from above at normal and oblique incidence and through the glass, with the
  tables' background flux closing the total-field box's balance (float32 here: bounds stated per quantity);
* a lossless dielectric sphere on glass absorbs nothing (the box balance with the layered background closes to a
  fraction of its extinction) while it scatters.
"""

import numpy as np
import pytest

from lightsmith_amd import api
from lightsmith_amd import problem as P
from lightsmith_amd.fdtd import scattering, workloads
from tests import cases
from tests.test_gpu_au_sphere import _advance_source, _rel
from tests.test_substrate import GLASS, sim_on

pytestmark = pytest.mark.gpu
AREA = np.pi * 20.0 ** 2  # the test sphere's geometric cross section, nm^2


def test_metal_sphere_on_glass_matches_the_fp64_oracle(gpu):
    au = P.DrudeLorentz(**workloads.AU_DRUDE_2LORENTZ)
    sim = sim_on(GLASS, direction=(0.643, 0.0, -0.766), polarization=(0.766, 0.0, 0.643), material=au, z_c=20.5)
    opts = P.FDTDOptions.model_validate(P.options_for(sim, "lightsmith-fdtd", "fdtd"))
    scene = scattering.setup(sim, opts, gpu)
    assert any(hasattr(g, "slot_vert") for g in scene.groups)
    _advance_source(scene.source, int(0.8 * scene.source.t_passed / scene.dt))
    g = scene.engine(gpu)
    o = cases.oracle_scattering_engine(scene)
    o32 = cases.oracle_scattering_engine(scene, dtype=np.float32)
    for q in range(3):
        for e in (g, o, o32):
            e.chunk()
        print(f"chunk {q + 1}/3 ({scene.stride} steps)", flush=True)

    def errors(a):
        Ea, Ha = cases.fields(a)
        Eo, Ho = cases.fields(o)
        out = {"E": _rel(Ea, Eo), "H": _rel(Ha, Ho)}
        for name, ma, mo in zip(("sca", "tot", "inc"), a.monitors, o.monitors):
            ea, ha = ma.results()
            eo, ho = mo.results()
            out[name + "_e"] = _rel(ea, eo)
            if ho is not None:
                out[name + "_h"] = _rel(ha, ho)
        return out

    errs, floor = errors(g), errors(o32)
    print(f"Au sphere on glass {scene.layout.shape}, {scene.stats}: " +
          ", ".join(f"{k} {v:.2e} (float32 floor {floor[k]:.2e})" for k, v in errs.items()))
    g.close()
    assert all(v < max(1e-5, 1.25 * floor[k]) for k, v in errs.items()), errs


@pytest.mark.parametrize("direction, polarization, label", [
    ((0.0, 0.0, -1.0), (1.0, 0.0, 0.0), "normal, from above"),
    ((0.643, 0.0, -0.766), (0.0, 1.0, 0.0), "40 deg s, from above"),
    ((0.5, 0.0, 0.866), (0.866, 0.0, -0.5), "30 deg p, through the glass"),
])
def test_null_particle_on_glass_scatters_and_absorbs_nothing(gpu, direction, polarization, label):
    sim = sim_on(GLASS, direction=direction, polarization=polarization, material=P.ConstantIndex(n_real=1.0),
                 z_c=26.0)  # 6 nm (over two cells) clear of the interface, whose cells keep the tables' meshing
    res = api.solve(sim, "fdtd", gpu)
    cs, ca = np.abs(res.c_sca_nm2).max(), np.abs(res.c_abs_nm2).max()
    print(f"null particle on glass, {label}: |C_sca| {cs:.2e}, |C_abs| {ca:.2e} nm^2 (pi r^2 = {AREA:.0f}); "
          f"steps {res.provenance['steps']}, {res.provenance['substrate']}")
    assert res.provenance["substrate"]["source"] == "layered tables"
    assert cs < 1e-6 * AREA and ca < 2e-4 * AREA


def test_lossless_sphere_on_glass_absorbs_nothing(gpu):
    res = api.solve(sim_on(GLASS, direction=(0.643, 0.0, -0.766), polarization=(0.0, 1.0, 0.0)), "fdtd", gpu)
    ce, ca, cs = (np.asarray(v) for v in (res.c_ext_nm2, res.c_abs_nm2, res.c_sca_nm2))
    print(f"n = 2 sphere on glass, 40 deg s: C_sca {cs}, C_abs {ca}, C_ext {ce}")
    assert np.all(cs > 0.5) and np.abs(ca).max() < 1e-3 * np.abs(ce).max()  # measured 6e-5 of C_ext


def _oblique(theta, pol, below):
    """(direction, polarisation) of a plane wave in the xz plane at theta from the normal of its incidence medium:
    s has E along y, p in the plane of incidence."""
    t = np.radians(theta)
    k = np.array([np.sin(t), 0.0, np.cos(t) if below else -np.cos(t)])
    e = np.array([0.0, 1.0, 0.0]) if pol == "s" else np.cross([0.0, 1.0, 0.0], k)
    return tuple(k), tuple(e / np.linalg.norm(e))


@pytest.mark.parametrize("theta, pol, below, rtol", [
    (0.0, "s", False, 5e-3), (45.0, "s", False, 5e-3), (45.0, "p", False, 8e-3),
    (0.0, "s", True, 5e-3), (30.0, "p", True, 5e-3),
])
def test_sphere_above_glass_matches_the_dipole_model(gpu, theta, pol, below, rtol):
    from oracle.dipole_substrate import sphere_above_substrate
    wl = [400.0, 500.0, 600.0, 800.0]
    k, e = _oblique(theta, pol, below)
    sim = sim_on(GLASS, direction=k, polarization=e, z_c=40.0, spacing_nm=2.0, pml_cells=16).model_copy(
        update={"wavelengths": wl})
    res = api.solve(sim, "fdtd", gpu)
    ref = sphere_above_substrate(20.0, 2.0, 1.0, 1.5, 40.0, wl, theta, pol, from_below=below)
    rel = np.abs(np.asarray(res.c_sca_nm2) / ref - 1)
    print(f"sphere 40 nm above glass, {theta} deg {pol}{' through the glass' if below else ''}: C_sca "
          f"{np.round(res.c_sca_nm2, 4)} vs dipole model {np.round(ref, 4)}: rel {rel}; grid {res.provenance['grid']}")
    assert np.all(rel < rtol)
