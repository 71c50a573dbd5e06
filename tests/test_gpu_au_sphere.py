"""The plain Au sphere end to end on the MI355X (r = 40 nm Drude + 2 Lorentz gold in water, 97^3 nodes at 2 nm with a
12-cell CPML, table-fed plane wave over 400-800 nm, 81-wavelength flux boxes):

* at the full grid, 3 DFT chunks (321 steps, the source started 3200 steps into its pulse) of the HIP engine against
  the FP64 oracle on the same inputs (the product's set-up: conformal voxelisation with the vertex tensor, source
  tables, flux boxes): fields and every monitor spectrum within 3e-5 relative L2. This is a stated deviation from
  the 1e-5 target: the reference's algorithm itself run in float32 (the oracle with dtype float32) is 2.1e-5 (E),
  7.8e-6 (H), 1.3e-5 (scattered-field E spectra) from float64 on this case, and the HIP engine is 2.0e-5, 7.6e-6,
  1.3e-5: the float32 floor of a Drude metal under an abrupt switch-on, not an implementation error;
* 4000 steps (DFT every 100) against the oracle's committed fixture (tests/golden/au_sphere_4000.npz): the incident
  reference, both boxes' net fluxes and 64 seeded points of each box's spectra within 1e-5, or within 1.25 times the
  float32 floor the fixture records for that quantity where the floor is higher (a stated deviation: the
  scattered-field flux of the float32 reference algorithm is 1.1e-4 from float64, its box spectra 3-6e-5, mostly
  from its float32 source tables; the HIP engine measured the same to two digits);
* the decay-terminated solve through ``api.solve(sim, "fdtd", "gpu")`` against Mie (oracle/mie.py) with the same
  eps(omega), at the reference's bounds for its gold sphere (tests/test_fdtd.py:446-456): peak within 5 nm, peak
  height within 3 %, C_ext within 7 % over that test's band (480-620 nm); C_abs > 0. Over the whole 400-800 nm band
  the 2 nm grid is within 10 %: the error grows towards 800 nm (9 % there), where it halves on a 1 nm grid (4 %),
  the first-order convergence of the metal interface, not a defect of the run (a 1e-8 decay threshold changes
  nothing).
"""

import os

import numpy as np
import pytest
import torch

from tests import cases

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
TOL = 1e-5
TOL_F32_FLOOR = 3e-5  # the full-grid metal case, whose float32 floor is 2.1e-5 (see the module docstring)


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


def _advance_source(src, steps: int):
    """Start the table-fed wave `steps` steps into its pulse (both engines read the same shifted tables): from rest,
    the box then switches on mid-pulse, so a short run already carries the full incident and scattered fields (the
    pulse centre otherwise arrives after ~3300 steps)."""
    for kind, (nodes, base, i0, u0, coef) in list(src.points.items()):
        src.points[kind] = (nodes, base, i0 + steps, u0, coef)
    base, i0, u0 = src.inc
    src.inc = (base, i0 + steps, u0)


def test_au_sphere_full_grid_matches_the_fp64_oracle(gpu):
    from lightsmith_amd.fdtd import workloads
    sim, opts, scene = workloads.au_sphere_scene(gpu)
    _advance_source(scene.source, 3200)
    g = scene.engine(gpu)
    assert g.fused_table_tensor  # the checked path is the one the solve runs
    o = cases.oracle_scattering_engine(scene)
    chunks = 3
    for q in range(chunks):
        g.chunk()
        o.chunk()
        print(f"chunk {q + 1}/{chunks} ({scene.stride} steps)", flush=True)
    Eg, Hg = cases.fields(g)
    Eo, Ho = cases.fields(o)
    errs = {"E": _rel(Eg, Eo), "H": _rel(Hg, Ho)}
    for name, mg, mo in zip(("sca", "tot", "inc"), g.monitors, o.monitors):
        eg, hg = mg.results()
        eo, ho = mo.results()
        errs[name + "_e"] = _rel(eg, eo)
        if ho is not None:
            errs[name + "_h"] = _rel(hg, ho)
    print(f"Au sphere {scene.layout.shape}, {chunks * scene.stride} steps from 3200 steps into the pulse, "
          f"{scene.stats}: |E| {np.linalg.norm(Eo):.3e}: " +
          ", ".join(f"{k} {v:.2e}" for k, v in errs.items()))
    g.close()
    assert all(v < TOL_F32_FLOOR for v in errs.values()), errs


def test_fused_table_and_tensor_launch_is_bit_identical(gpu):
    """The tensor nodes' D update launched with the E-side table correction (one kernel) against the two launches:
    every state tensor equal after 3 chunks, with the source 3200 steps into its pulse."""
    from lightsmith_amd.fdtd import workloads
    from lightsmith_amd.fdtd.engine import Engine
    sim, opts, scene = workloads.au_sphere_scene(gpu)
    _advance_source(scene.source, 3200)
    fused = scene.engine(gpu)
    Engine.fuse_sparse = False
    try:
        alone = scene.engine(gpu)
    finally:
        Engine.fuse_sparse = True
    assert fused.fused_table_tensor and not alone.fused_table_tensor
    for _ in range(3):
        fused.chunk()
        alone.chunk()
    same = [torch.equal(a, b) for a, b in zip(fused.state_tensors(), alone.state_tensors())]
    print(f"fused vs separate launches: {sum(same)}/{len(same)} state tensors equal", flush=True)
    fused.close()
    alone.close()
    assert all(same)


def test_au_sphere_4000_steps_match_the_oracle_fixture(gpu):
    from lightsmith_amd.fdtd import scattering, workloads
    from lightsmith_amd.fdtd.run import flux
    ref = np.load(os.path.join(HERE, "golden", "au_sphere_4000.npz"))
    steps, stride = int(ref["steps"]), int(ref["stride"])
    sim, opts, scene = workloads.au_sphere_scene(gpu, dft_stride=stride)
    g = scene.engine(gpu)
    run, used = g.make_runner(True)
    for _ in range(steps // stride):
        run()
    g.release()
    (es, hs), (et, ht), (e_ref, _) = (m.results() for m in g.monitors)
    c_sca, c_abs, c_ext = scattering.cross_sections(scene, e_ref, es, hs, et, ht)
    errs = {"e_ref": _rel(e_ref, ref["e_ref"]), "flux_sca": _rel(flux(es, hs, scene.sca[3]), ref["flux_sca"]),
            "flux_tot": _rel(flux(et, ht, scene.tot[3]), ref["flux_tot"]),
            "es": _rel(es[:, ref["pick_sca"]], ref["es"]), "hs": _rel(hs[:, ref["pick_sca"]], ref["hs"]),
            "et": _rel(et[:, ref["pick_tot"]], ref["et"]), "ht": _rel(ht[:, ref["pick_tot"]], ref["ht"])}
    errs["c_ext"] = _rel(c_ext, ref["c_ext"])
    tol = {k: max(TOL, 1.25 * float(ref["f32_" + k])) for k in errs}
    print(f"Au sphere, {steps} steps (graphs {used}): " +
          ", ".join(f"{k} {v:.2e} (float32 floor {float(ref['f32_' + k]):.2e})" for k, v in errs.items()))
    g.close()
    assert all(errs[k] < tol[k] for k in errs), (errs, tol)


def test_au_sphere_solve_matches_mie(gpu):
    from lightsmith_amd import api
    from lightsmith_amd.fdtd import workloads
    from oracle.mie import mie_q
    sim = workloads.au_sphere_sim()
    res = api.solve(sim, method="fdtd", device="gpu")
    wl = np.asarray(res.wavelengths_nm)
    mat = sim.particles[0].material
    n_h = sim.medium_index
    r = sim.particles[0].shape.radius_nm
    q = np.array([mie_q(np.sqrt(complex(mat.eps(np.array([w]))[0])) / n_h, 2 * np.pi * n_h * r / w)[0] for w in wl])
    mie = q * np.pi * r ** 2
    fd = np.asarray(res.c_ext_nm2)
    i, j = int(np.argmax(mie)), int(np.argmax(fd))
    rel = np.abs(fd - mie) / mie
    print(f"Au sphere via api.solve: {res.provenance['steps']} steps, grid {res.provenance['grid']}, "
          f"{res.provenance['tensor_vertices']} tensor vertices; peak {wl[j]:.0f} nm (Mie {wl[i]:.0f} nm), height "
          f"{fd[j] / mie[i] - 1:+.2%}, max |C_ext - Mie| / Mie {rel.max():.2%}, decay {res.provenance['decay_level']:.1e}")
    assert res.backend == "lightsmith-fdtd" and res.provenance["hip_graphs"]
    band = (wl >= 480) & (wl <= 620)
    print(f"max |C_ext - Mie| / Mie over 480-620 nm {rel[band].max():.2%}")
    assert abs(wl[i] - wl[j]) <= 5.0
    assert abs(fd[j] / mie[i] - 1) < 0.03
    assert rel[band].max() < 0.07
    assert rel.max() < 0.10
    assert min(res.c_abs_nm2) > 0
