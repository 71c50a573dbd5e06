"""Layer stacks and periodic structures solved end to end on the GPU through the backend (lightsmith_amd/fdtd/
spectra.py), against the characteristic-matrix TMM (oracle/tmm.py) with the reference's tolerances
(test_fdtd_periodic.py:1-20, 46-82), and against what symmetry and energy fix for patterned cells: a lossless
grating or pillar array conserves energy, a mirror-symmetric grating at normal incidence sends equal power into
orders +m and -m, a C4-symmetric pillar array reflects s and p alike at normal incidence. The batch of runs
(one launch per half-step over every run) equals the runs stepped one after another bit for bit."""

import math

import numpy as np
import pytest

from lightsmith_amd import api
from lightsmith_amd import problem as P
from lightsmith_amd.fdtd import spectra
from oracle.tmm import stack_rt
from tests.test_spectra import lamellar, pillars, stack

pytestmark = pytest.mark.gpu
CI = P.ConstantIndex


def _tmm(prob):
    """R, T from the oracle per polarisation as [angle][wavelength] arrays."""
    wl = prob.wavelength_array()
    eps_l = [lay.material.eps(wl) for lay in prob.layers]
    d = [lay.thickness_nm for lay in prob.layers]
    out = {}
    for pol in prob.polarizations:
        rt = [stack_rt(eps_l, d, prob.superstrate.eps(wl), prob.substrate.eps(wl), wl, math.radians(a), pol)
              for a in prob.angles_deg]
        out[pol] = (np.array([q[0] for q in rt]), np.array([q[1] for q in rt]))
    return out


def _max_diff(res, ref):
    return max(max(np.abs(np.asarray(res.r[p]) - ref[p][0]).max(), np.abs(np.asarray(res.t[p]) - ref[p][1]).max())
               for p in ref)


def test_dielectric_slab_and_drude_film_match_tmm(gpu):
    """n = 2 slab 300 nm on glass, 400-800 nm, 10 nm cells: tolerance 0.01 (the reference measures 0.006);
    Drude film 30 nm, s and p: 0.005 (reference 0.0025); at normal incidence one run serves both polarisations."""
    for prob, tol in ((stack(CI(n_real=2.0), 300), 0.01),
                      (stack(P.Drude(eps_inf=5, wp_ev=9.0, gamma_ev=0.05), 30, pols=("s", "p")), 0.005)):
        res = api.solve(prob, "fdtd", gpu)
        d = _max_diff(res, _tmm(prob))
        print(f"{prob.layers[0].material.kind} {prob.layers[0].thickness_nm} nm: max |dR|, |dT| vs TMM {d:.4f}, "
              f"runs {len(res.provenance['runs'])}, steps {res.provenance['steps']}, "
              f"{res.provenance['time_loop_s']:.2f} s")
        assert res.kind == "spectrum" and res.provenance["precision"] == "single"
        assert res.provenance["real_cells"][:2] == [1, 1] and len(res.provenance["runs"]) == 1
        assert d < tol
    assert res.r["s"] == res.r["p"]


def test_oblique_slab_matches_tmm_and_batch_equals_sequential(gpu):
    """n = 2 slab 300 nm at 0, 30 and 60 deg, s and p, 450 and 750 nm: tolerance 0.008 (the reference measures
    <= 0.0046); one broadband run at normal incidence, one complex fixed-k_par run per (angle, wavelength,
    polarisation) off-normal, 9 runs, stepped as one batch and then one after another."""
    prob = P.LayerStackProblem(superstrate=CI(n_real=1.0), substrate=CI(n_real=1.5),
                               layers=[P.Layer(thickness_nm=300, material=CI(n_real=2.0))], wavelengths=[450.0, 750.0],
                               angles_deg=[0.0, 30.0, 60.0], polarizations=["s", "p"])
    opts = P.FDTDOptions()
    res = spectra.solve(prob, opts, gpu)
    ref = _tmm(prob)
    d = _max_diff(res, ref)
    print(f"oblique slab: max |dR|, |dT| vs TMM {d:.4f}; steps per run "
          f"{[q['steps'] for q in res.provenance['runs']]}, {res.provenance['time_loop_s']:.2f} s batched")
    assert res.angles_deg == [0.0, 30.0, 60.0] and len(res.provenance["runs"]) == 1 + 2 * 2 * 2
    assert d < 0.008
    assert np.abs(np.asarray(res.a["s"])).max() < 2e-3 and not res.warnings
    assert not np.allclose(res.r["s"][2], res.r["p"][2])
    assert "fixed-k_par" in res.provenance["notes"] and res.provenance["batched"]
    seq = spectra.solve(prob, opts, gpu, batched=False)
    print(f"sequential: {seq.provenance['time_loop_s']:.2f} s")
    assert seq.r == res.r and seq.t == res.t
    assert [q["steps"] for q in seq.provenance["runs"]] == [q["steps"] for q in res.provenance["runs"]]


def test_lamellar_grating_conserves_energy_and_is_mirror_symmetric(gpu):
    """P = 700 nm, 300 nm eps 6 lines, 250 nm tall, on glass, s and p at normal incidence, away from the Rayleigh
    anomaly at 700 nm: R + T = 1 within 0.01 (lossless), orders +m and -m carry equal power (the cell is mirror
    symmetric about x = 0 on the grid too), the reported orders are the propagating ones."""
    wl = [500.0, 600.0, 800.0, 900.0]
    res = api.solve(lamellar(wl), "fdtd", gpu)
    a = max(np.abs(np.asarray(res.a[p])).max() for p in "sp")
    orders = {o.order: o for o in res.orders}
    asym = max(np.abs(np.asarray(orders[(m, 0)].r[p]) - np.asarray(orders[(-m, 0)].r[p])).max()
               + np.abs(np.asarray(orders[(m, 0)].t[p]) - np.asarray(orders[(-m, 0)].t[p])).max()
               for m in (1, 2) for p in "sp")
    print(f"lamellar: |A| max {a:.2e}, +-m asymmetry {asym:.1e}, orders {sorted(orders)}; "
          f"R_s {np.round(res.r['s'][0], 4)}, T_s {np.round(res.t['s'][0], 4)}")
    assert sorted(orders) == [(-2, 0), (-1, 0), (0, 0), (1, 0), (2, 0)]
    assert a < 0.01 and asym < 1e-3
    # orders +-2 propagate in the substrate only (2 x 500 / 700 lies between 1 and 1.5): nothing reflected into them
    assert np.all(np.asarray(orders[(2, 0)].r["s"]) == 0.0)
    t1 = np.asarray(orders[(1, 0)].t["s"][0])
    assert np.all(t1 > 1e-3)  # the first orders carry power at every wavelength


def test_pillar_array_conserves_energy_is_polarisation_independent_and_batch_equals_sequential(gpu):
    """Square lattice (400 nm) of n = 2 disks (r 120 nm, 150 nm) on n = 1.45 at normal incidence: lossless, so
    R + T = 1 within 0.01; the cell is C4 symmetric on the grid (equal cells along x and y), so s and p reflect
    alike; the batch of the two runs equals the runs one after another bit for bit."""
    prob = pillars()
    opts = P.FDTDOptions()
    res = spectra.solve(prob, opts, gpu)
    a = max(np.abs(np.asarray(res.a[p])).max() for p in "sp")
    ds = max(np.abs(np.asarray(res.r["s"]) - np.asarray(res.r["p"])).max(),
             np.abs(np.asarray(res.t["s"]) - np.asarray(res.t["p"])).max())
    print(f"pillars: grid {res.provenance['grid']}, |A| max {a:.2e}, |R_s - R_p| {ds:.1e}, "
          f"R_s {np.round(res.r['s'][0], 4)}, orders {[o.order for o in res.orders]}")
    assert a < 0.01 and ds < 1e-3
    assert res.provenance["offdiagonal_cells"] > 0  # the disks' conformal interfaces
    seq = spectra.solve(prob, opts, gpu, batched=False)
    assert seq.r == res.r and seq.t == res.t


def test_uniform_periodic_cell_is_the_layer_stack(gpu):
    """A periodic cell whose layer has no shapes is a planar stack: its zeroth orders match TMM like the stack's,
    and no other order carries power."""
    wl = [450.0, 600.0, 750.0]
    prob = P.PeriodicProblem(lattice_a1_nm=(300, 0), lattice_a2_nm=(0, 300), superstrate=CI(n_real=1.0),
                             substrate=CI(n_real=1.5),
                             layers=[P.PatternedLayer(thickness_nm=200, background=CI(n_real=2.0))], wavelengths=wl,
                             incidence=P.Incidence(polarizations=["s"]))
    res = api.solve(prob, "fdtd", gpu)
    st = P.LayerStackProblem(superstrate=CI(n_real=1.0), substrate=CI(n_real=1.5),
                             layers=[P.Layer(thickness_nm=200, material=CI(n_real=2.0))], wavelengths=wl,
                             polarizations=["s"])
    ref = _tmm(st)
    d = _max_diff(res, ref)
    other = max([np.abs(np.asarray(o.r["s"])).max() + np.abs(np.asarray(o.t["s"])).max()
                 for o in res.orders if o.order != (0, 0)], default=0.0)
    print(f"uniform periodic cell: max |dR|, |dT| vs TMM {d:.4f}; other orders {other:.1e}")
    assert d < 0.01 and other < 1e-4


def _rcwa_orders(prob, theta_deg, n_harm=30):
    """Per polarisation {order m: (R_m, T_m) per wavelength} of a one-layer lamellar problem from the oracle."""
    from oracle.rcwa import lamellar as rcwa
    lay = prob.layers[0]
    boxes = [(s.profile.center_nm[0], s.profile.size_nm[0], s.material.eps([600.0])[0]) for s in lay.shapes]
    layers = [(lay.thickness_nm, lay.background.eps([600.0])[0], boxes)]
    out = {}
    for pol in prob.incidence.polarizations:
        rows = [rcwa(layers, 1.0, 2.25, prob.lattice_a1_nm[0], lam, math.radians(theta_deg), pol, n_harm)
                for lam in prob.wavelength_array()]
        n = rows[0][0]
        out[pol] = {int(m): (np.array([r[1][i] for r in rows]), np.array([r[2][i] for r in rows]))
                    for i, m in enumerate(n)}
    return out


@pytest.mark.parametrize("theta, wl, tol, tol_order", [
    (0.0, [500.0, 600.0, 800.0, 900.0, 1000.0, 1200.0], 0.015, 0.01),  # away from the anomaly at 700 nm
    (30.0, [500.0, 800.0, 1200.0], 0.008, 0.012),  # Rayleigh anomalies at 700 and 1050 nm
])
def test_lamellar_grating_matches_rcwa(gpu, theta, wl, tol, tol_order):
    """The reference's lamellar grating (P = 700 nm, 300 nm lines of eps 6, 250 nm, on glass) against the RCWA
    oracle (30 harmonics, Li's rule for p), with the reference's tolerances (test_fdtd_periodic.py:145-192: totals
    0.015 at normal incidence, 0.008 at 30 deg; single orders 0.01 and 0.012). Every order the oracle finds carrying
    power is reported, and each reported order matches."""
    prob = lamellar(wl, theta=theta)
    res = api.solve(prob, "fdtd", gpu)
    ref = _rcwa_orders(prob, theta)
    worst, worst_order = 0.0, 0.0
    for pol in "sp":
        r_tot = sum(v[0] for v in ref[pol].values())
        t_tot = sum(v[1] for v in ref[pol].values())
        worst = max(worst, np.abs(np.asarray(res.r[pol][0]) - r_tot).max(),
                    np.abs(np.asarray(res.t[pol][0]) - t_tot).max())
        mine = {o.order[0]: o for o in res.orders}
        for m, (rm, tm) in ref[pol].items():
            if m not in mine:
                assert rm.max() + tm.max() < 1e-6, (pol, m)
                continue
            worst_order = max(worst_order, np.abs(np.asarray(mine[m].r[pol][0]) - rm).max(),
                              np.abs(np.asarray(mine[m].t[pol][0]) - tm).max())
    print(f"lamellar {theta} deg vs RCWA: totals {worst:.4f}, orders {worst_order:.4f}; "
          f"runs {len(res.provenance['runs'])}, {res.provenance['time_loop_s']:.2f} s")
    assert worst < tol and worst_order < tol_order


@pytest.mark.parametrize("substrate, lam0", [
    (P.Drude(eps_inf=9.5, wp_ev=9.0, gamma_ev=0.07), 480.0),  # a Drude metal, eps -4.1 + 0.3i .. -13.6 + 0.8i
    (P.DrudeLorentz(eps_inf=3.0, poles=[P.LorentzPole(delta_eps=1.5, omega0_ev=3.0, gamma_ev=1.5)]), 420.0),  # ~4.5 + 1i
], ids=["metal", "lossy_dielectric"])
def test_absorbing_substrate_transmission_is_taken_at_the_interface(gpu, substrate, lam0):
    """80 nm of n = 1.46 on an absorbing substrate, up to 700 nm, at 0 and 30 deg: T is the flux into the substrate
    at the interface, as TMM gives it (the monitor's depth undone per order), so a metal substrate does not report
    T = 0. Tolerance 0.02, the reference's for its fitted Si and Au substrates (test_fdtd_periodic.py:268-279). The
    metal's band starts at 480 nm, clear of its eps = 0 point near 420 nm (420-700 nm: 0.021)."""
    prob = P.LayerStackProblem(superstrate=CI(n_real=1.0), layers=[P.Layer(thickness_nm=80, material=CI(n_real=1.46))],
                               substrate=substrate, wavelengths=P.WavelengthRange(start_nm=lam0, stop_nm=700, num=5),
                               angles_deg=[0.0, 30.0], polarizations=["s", "p"])
    res = api.solve(prob, "fdtd", gpu)
    ref = _tmm(prob)
    d = _max_diff(res, ref)
    t = np.asarray(res.t["s"])
    per_wl = np.max([np.abs(np.asarray(getattr(res, q)[p]) - ref[p][i]).max(axis=0)
                     for p in ref for i, q in enumerate("rt")], axis=0)
    print(f"80 nm film on {substrate.kind}: max |dR|, |dT| vs TMM {d:.4f} (per wavelength {np.round(per_wl, 4)}); "
          f"T_s {np.round(t[0], 4)}")
    assert d < 0.02 and t.min() > 1e-3
