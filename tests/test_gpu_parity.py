"""HIP engine vs the FP64 oracle on the same inputs (tests/cases.py).

Tolerance: relative L2 error of the float32 GPU fields against the float64 oracle after the stated number
of steps: FIELD_TOL = 1e-5, or -- where a configuration is so conditioned that float32 arithmetic alone moves
the fields further -- twice the error of the reference algorithm run in float32 (the oracle with
dtype=float32, the reference's single-precision op order), whichever is larger. DFT spectra: DFT_TOL, same
rule. Each test prints the measured errors.
"""

import numpy as np
import pytest
import torch

from tests import cases

pytestmark = pytest.mark.gpu

FIELD_TOL = 1e-5
DFT_TOL = 1e-5


def _compare(o, g, label, mask=None, tol=FIELD_TOL, o32=None):
    Eo, Ho = cases.fields(o)
    Eg, Hg = cases.fields(g)
    eE, eH = cases.rel_l2(Eg, Eo, mask), cases.rel_l2(Hg, Ho, mask)
    extra = ""
    if o32 is not None:
        E3, H3 = cases.fields(o32)
        fE, fH = cases.rel_l2(E3, Eo, mask), cases.rel_l2(H3, Ho, mask)
        tol = max(tol, 2 * max(fE, fH))
        extra = f"  (reference algorithm in float32: E {fE:.2e} H {fH:.2e}; tolerance {tol:.2e})"
    print(f"{label}: rel L2 E {eE:.2e}  H {eH:.2e}{extra}")
    assert eE < tol and eH < tol, (eE, eH, tol)
    return eE, eH


def test_pec_box_random_fields(gpu):
    o = cases.make(cases.oracle_ns(), shape=(70, 33, 29), dx=1.0, dt=0.95 / np.sqrt(3),
                   inv_eps=np.full((3, 70, 33, 29), 1 / 2.25))
    g = cases.make(cases.product_ns(), gpu, shape=(70, 33, 29), dx=1.0, dt=0.95 / np.sqrt(3),
                   inv_eps=np.full((3, 70, 33, 29), 1 / 2.25))
    cases.random_fields(o, 1)
    cases.random_fields(g, 1)
    o.run(60)
    g.steps(60)
    _compare(o, g, "PEC box 60 steps")


@pytest.mark.parametrize("tiling", [(1, 4, 64, 5), (2, 4, 64, 16), (2, 8, 64, 7), (1, 8, 64, 3), (1, 4, 32, 6),
                                    (2, 8, 32, 4)])
def test_tilings_agree(gpu, tiling):
    shape = (67, 45, 23)
    eps = np.random.default_rng(3).uniform(1, 4, size=(3,) + shape)
    o = cases.make(cases.oracle_ns(), shape=shape, dx=1.0, dt=0.9 / np.sqrt(3), inv_eps=1 / eps,
                   pml=cases.oracle_ns().PMLSpec(cells=((5, 4), (3, 6), (4, 4)), alpha_max=0.02))
    ns = cases.product_ns()
    old = ns.Engine.tiling
    ns.Engine.tiling = tiling
    try:
        g = cases.make(ns, gpu, shape=shape, dx=1.0, dt=0.9 / np.sqrt(3), inv_eps=1 / eps,
                       pml=ns.PMLSpec(cells=((5, 4), (3, 6), (4, 4)), alpha_max=0.02))
    finally:
        ns.Engine.tiling = old
    cases.random_fields(o, 2)
    cases.random_fields(g, 2)
    o.run(40)
    g.steps(40)
    _compare(o, g, f"tiling {tiling}")


@pytest.mark.parametrize("material", ["indexed", "dense"])
def test_dispersive_cpml_dipole_box(gpu, material):
    steps = 300
    o = cases.dispersive_box(cases.oracle_ns(), n_steps_hint=steps)
    g = cases.dispersive_box(cases.product_ns(), gpu, n_steps_hint=steps, material=material)
    assert g.material_mode == material
    assert g.n_ade > 0
    for _ in range(steps // o.stride):
        o.chunk()
        g.chunk()
    _compare(o, g, f"dispersive box ({material}) {steps} steps")
    eo, ho = o.monitors[0].results()
    eg, hg = g.monitors[0].results()
    de, dh = cases.rel_l2(eg, eo), cases.rel_l2(hg, ho)
    print(f"DFT rel L2 E {de:.2e} H {dh:.2e}; energy {float(g.energy[0]):.6e} vs {o.energy[0]:.6e}")
    assert de < DFT_TOL and dh < DFT_TOL
    assert abs(float(g.energy[0]) / o.energy[0] - 1) < 1e-4


def test_cpml_kappa_alpha(gpu):
    steps = 200
    o = cases.dispersive_box(cases.oracle_ns(), n_steps_hint=steps, kappa_max=3.0, alpha=0.05, frac=0.0)
    g = cases.dispersive_box(cases.product_ns(), gpu, n_steps_hint=steps, kappa_max=3.0, alpha=0.05, frac=0.0)
    for _ in range(steps // o.stride):
        o.chunk()
        g.chunk()
    _compare(o, g, "CPML kappa/alpha")


def test_bloch_periodic_complex(gpu):
    shape = (20, 14, 40)
    dx = (3.0, 2.5, 2.0)
    dt = 0.9 / np.sqrt(sum(1 / d ** 2 for d in dx))
    kpar = (0.05, -0.03, 0.0)
    eps = np.random.default_rng(7).uniform(1, 6, size=(3,) + shape)
    kw = dict(shape=shape, dx=dx, dt=dt, inv_eps=1 / eps, periodic=(True, True, False), bloch=kpar)
    on, pn = cases.oracle_ns(), cases.product_ns()
    o = cases.make(on, pml=on.PMLSpec(cells=((0, 0), (0, 0), (6, 6))), **kw)
    g = cases.make(pn, gpu, pml=pn.PMLSpec(cells=((0, 0), (0, 0), (6, 6))), **kw)
    rng = np.random.default_rng(5)
    for c in range(3):
        for F, R in (("E", o.RE[c]), ("H", o.RH[c])):
            vals = rng.standard_normal(tuple(r.stop - r.start for r in R)) + 1j * rng.standard_normal(
                tuple(r.stop - r.start for r in R))
            getattr(o, F)[c][R] = vals
            getattr(g, F)[c][R] = torch.as_tensor(vals, dtype=torch.complex64, device=g.E_flat.device)
    o.run(80)
    g.steps(80)
    _compare(o, g, "Bloch periodic complex", mask=cases.real_mask(shape, (True, True, False)))


def test_graph_replay_matches_eager(gpu):
    steps = 200
    a = cases.dispersive_box(cases.product_ns(), gpu, n_steps_hint=steps)
    b = cases.dispersive_box(cases.product_ns(), gpu, n_steps_hint=steps)
    run, used = b.make_runner(True)
    assert used
    for _ in range(steps // a.stride):
        a.chunk()
        run()
    b.release()
    Ea, Ha = cases.fields(a)
    Eb, Hb = cases.fields(b)
    assert np.array_equal(Ea, Eb) and np.array_equal(Ha, Hb)
    ea, _ = a.monitors[0].results()
    eb, _ = b.monitors[0].results()
    assert np.array_equal(ea, eb)
    assert int(a.n[0]) == int(b.n[0]) == steps


@pytest.mark.parametrize("khat", [(0, 0, 1.0), (1.0, 2.0, 3.0)])
def test_tfsf_plane_wave_box(gpu, khat):
    steps = 300
    k = np.asarray(khat, float) / np.linalg.norm(khat)
    e = np.cross(k, [0.0, 1.0, 0.0]) if abs(k[1]) < 0.9 else np.cross(k, [1.0, 0, 0])
    e /= np.linalg.norm(e)
    o = cases.tfsf_box(cases.oracle_ns(), khat=khat, ehat=e, steps=steps)
    g = cases.tfsf_box(cases.product_ns(), gpu, khat=khat, ehat=e, steps=steps)
    for _ in range(steps // o.stride):
        o.chunk()
        g.chunk()
    _compare(o, g, f"1D-fed TFSF {khat}")
    eo, ho = o.monitors[0].results()
    eg, hg = g.monitors[0].results()
    assert cases.rel_l2(eg, eo) < DFT_TOL and cases.rel_l2(hg, ho) < DFT_TOL


def test_table_source(gpu):
    shape = (28, 27, 32)
    steps = 120
    on, pn = cases.oracle_ns(), cases.product_ns()
    inv = np.full((3,) + shape, 1 / 2.0)
    so, sp = cases.synthetic_table_source(on, shape), cases.synthetic_table_source(pn, shape)
    mo = on.DFTMonitor(np.array([0.3, 0.7]), np.arange(3), source=so)
    mp = pn.DFTMonitor(np.array([0.3, 0.7]), np.arange(3), source=sp)
    kw = dict(shape=shape, dx=1.0, dt=0.5, inv_eps=inv, dft_stride=5)
    o = cases.make(on, pml=on.PMLSpec(cells=((4, 4),) * 3), sources=[so], monitors=[mo], **kw)
    g = cases.make(pn, gpu, pml=pn.PMLSpec(cells=((4, 4),) * 3), sources=[sp], monitors=[mp], **kw)
    for _ in range(steps // 5):
        o.chunk()
        g.chunk()
    _compare(o, g, "table-fed TFSF")
    eo, _ = o.monitors[0].results()
    eg, _ = g.monitors[0].results()
    assert cases.rel_l2(eg, eo) < DFT_TOL


def test_bloch_plane_source(gpu):
    nz = 60
    shape = (11, 9, nz)
    sp = (20.0, 24.0, 15.0)
    kpar = (0.004, 0.002)
    dt = 0.9 / np.sqrt(sum(1 / d ** 2 for d in sp))
    steps = 400
    on, pn = cases.oracle_ns(), cases.product_ns()
    eps = np.full((3,) + shape, 1.0)
    eps[:, :, :, 35:45] = 2.1
    eps[:, 3:7, 2:6, 45:50] = 12.1
    out = []
    for ns, dev in ((on, None), (pn, gpu)):
        pml = ns.PMLSpec(cells=((0, 0), (0, 0), (8, 8)), eps_ref=1.0)
        src = ns.BlochPlaneSource(shape=shape, spacing=sp, dt=dt, k_src=25, k_launch=24, kpar=kpar, pol="p",
                                  eps_b=1.0, pml=pml, omega_c=2 * np.pi / 600, omega_half=0.3 * 2 * np.pi / 600,
                                  max_steps=steps + 10)
        i = np.arange(1, shape[0] - 1)
        j = np.arange(1, shape[1] - 1)
        gi, gj = (a.ravel() for a in np.meshgrid(i, j, indexing="ij"))
        kk = 52
        e = np.concatenate([ns.flat_index(0, gi, gj, kk, shape), ns.flat_index(1, gi, gj, kk, shape)])
        h1 = np.concatenate([ns.flat_index(1, gi, gj, kk - 1, shape), ns.flat_index(0, gi, gj, kk - 1, shape)])
        h2 = np.concatenate([ns.flat_index(1, gi, gj, kk, shape), ns.flat_index(0, gi, gj, kk, shape)])
        mons = [ns.DFTMonitor(np.array([2 * np.pi / 600]), e, h1, h2, energy=1),
                ns.DFTMonitor(np.array([2 * np.pi / 600]), *src.probe(30), source=src)]
        eng = cases.make(ns, dev, shape=shape, dx=sp, dt=dt, inv_eps=1 / eps, pml=pml, sources=[src], monitors=mons,
                         dft_stride=10, periodic=(True, True, False), bloch=(*kpar, 0.0))
        for _ in range(steps // 10):
            eng.chunk()
        out.append(eng)
    o, g = out
    _compare(o, g, "Bloch plane source", mask=cases.real_mask(shape, (True, True, False)))
    for mo, mg in zip(o.monitors, g.monitors):
        eo, ho = mo.results()
        eg, hg = mg.results()
        assert cases.rel_l2(eg, eo) < DFT_TOL and cases.rel_l2(hg, ho) < DFT_TOL


def test_mode_injection_and_time_monitor(gpu):
    shape = (24, 20, 26)
    rng = np.random.default_rng(11)
    on, pn = cases.oracle_ns(), cases.product_ns()
    jj, kk = np.meshgrid(np.arange(5, 15), np.arange(6, 20), indexing="ij")
    nwin = jj.size
    steps = 150
    r = 3
    tE = rng.standard_normal((steps + 5, r)) * np.exp(-((np.arange(steps + 5)[:, None] - 60) / 20.0) ** 2)
    tH = np.roll(tE, 1, axis=0)
    mH = rng.standard_normal((2 * nwin, r)) * 0.05
    mE = rng.standard_normal((2 * nwin, r)) * 0.05
    out = []
    for ns, dev in ((on, None), (pn, gpu)):
        h_idx = np.concatenate([ns.flat_index(1, 7, jj.ravel(), kk.ravel(), shape),
                                ns.flat_index(2, 7, jj.ravel(), kk.ravel(), shape)])
        e_idx = np.concatenate([ns.flat_index(1, 8, jj.ravel(), kk.ravel(), shape),
                                ns.flat_index(2, 8, jj.ravel(), kk.ravel(), shape)])
        src = ns.ModeInjection(h_idx, mH, tE, e_idx, mE, tH)
        tm = ns.TimeMonitor(e_idx[::7], 40)
        eng = cases.make(ns, dev, shape=shape, dx=1.0, dt=0.5, inv_eps=np.full((3,) + shape, 1 / 3.0),
                         pml=ns.PMLSpec(cells=((4, 4),) * 3), sources=[src], time_monitors=[tm], dft_stride=5)
        for _ in range(steps // 5):
            eng.chunk()
        out.append(eng)
    o, g = out
    _compare(o, g, "mode injection")
    assert cases.rel_l2(g.time_monitors[0].results(30), o.time_monitors[0].results(30)) < FIELD_TOL
    # past n_max samples the reference's indexed store raises; the engine reports the overflow at readback
    for _ in range(40 - steps // 5 + 1):
        g.chunk()
    with pytest.raises(IndexError, match="time monitor overflow"):
        g.time_monitors[0].results(40)


def test_mixed_groups_and_offdiagonal(gpu):
    shape = (26, 24, 22)
    rng = np.random.default_rng(4)
    on, pn = cases.oracle_ns(), cases.product_ns()
    out = []
    for ns, dev, dtype in ((on, None, np.float64), (on, None, np.float32), (pn, gpu, None)):
        metal = ns.Medium(2.5, (ns.Pole(0.02, 0.0, 0.01), ns.Pole(0.01, 0.12, 0.02)))
        diel = ns.Medium(4.0, ())
        host = ns.Medium(1.5, ())
        inv = np.full((3,) + shape, 1 / 1.5)
        # a mixed interface layer (two media: one dispersive) and a pure block
        ii, jj, kk = np.meshgrid(np.arange(8, 16), np.arange(8, 15), np.arange(9, 10), indexing="ij")
        idx_m = np.concatenate([ns.flat_index(c, ii.ravel(), jj.ravel(), kk.ravel(), shape) for c in range(3)])
        n = len(idx_m)
        frac = np.stack([np.full(n, 0.3), np.full(n, 0.7)])
        w = np.random.default_rng(1).uniform(0, 1, n)
        ii2, jj2, kk2 = np.meshgrid(np.arange(8, 16), np.arange(8, 15), np.arange(10, 14), indexing="ij")
        idx_p = np.concatenate([ns.flat_index(c, ii2.ravel(), jj2.ravel(), kk2.ravel(), shape) for c in range(3)])
        inv.reshape(-1)[idx_m] = 1.0
        inv.reshape(-1)[idx_p] = 1.0
        groups = [ns.DispersiveGroup(idx=idx_m, media=[host, metal], frac=frac, w=w),
                  ns.DispersiveGroup(idx=idx_p, media=[metal], frac=np.ones((1, len(idx_p))), w=None)]
        # off-diagonal terms on a dielectric slab elsewhere
        oi, oj, ok = np.meshgrid(np.arange(4, 20), np.arange(16, 19), np.arange(4, 18), indexing="ij")
        idx3 = np.stack([oi.ravel(), oj.ravel(), ok.ravel()], axis=1)
        idx3 = np.concatenate([idx3, idx3, idx3])
        comp = np.repeat(np.arange(3), len(oi.ravel()))
        coef = np.random.default_rng(2).uniform(-0.05, 0.05, (len(comp), 3))
        inv[:, 4:20, 16:19, 4:18] = 1 / 4.0
        od = ns.OffDiagonal(idx3=idx3, comp=comp, coef=coef)
        src = ns.PointDipole(shape=shape, origin=np.zeros(3), dx=1.0, dt=0.5, pos=(12.2, 11.1, 11.7), pdir=(1, 0.5, 0.2),
                             inv_eps=inv, omega_c=0.4, omega_half=0.15, max_steps=400)
        kw = {"dtype": dtype} if dtype is not None else {}
        eng = cases.make(ns, dev, shape=shape, dx=1.0, dt=0.5, inv_eps=inv, groups=groups, offdiag=od,
                         pml=ns.PMLSpec(cells=((4, 4),) * 3), sources=[src], dft_stride=10, **kw)
        for _ in range(30):
            eng.chunk()
        out.append(eng)
    o, o32, g = out
    _compare(o, g, "mixed groups + off-diagonal + pure", o32=o32)


@pytest.mark.parametrize("bloch", [False, True])
def test_vertex_tensor_sphere(gpu, bloch):
    """Conformal vertex tensor around a Drude sphere (TensorMesh), with pure D-form metal nodes, a dipole and a DFT
    plane; Bloch-periodic in x and y (complex64) in the second case."""
    kw = dict(shape=(26, 26, 24), periodic=(True, True, False), bloch=(0.05, 0.02, 0.0)) if bloch else {}
    steps = 300
    o = cases.tensor_sphere(cases.oracle_ns(), **kw)
    g = cases.tensor_sphere(cases.product_ns(), gpu, **kw)
    for _ in range(steps // o.stride):
        o.chunk()
        g.chunk()
    mask = cases.real_mask(o.shape, o.periodic) if bloch else None
    _compare(o, g, f"vertex tensor sphere ({'Bloch' if bloch else 'real'}) {steps} steps", mask)
    eo, _ = o.monitors[0].results()
    eg, _ = g.monitors[0].results()
    d = cases.rel_l2(eg, eo)
    print(f"  DFT rel L2 {d:.2e}")
    assert d < DFT_TOL


def _state_diff(a, b) -> tuple[float, float]:
    """(largest relative L2 difference over the state tensors, largest absolute difference)."""
    ta, tb = a.state_tensors(), b.state_tensors()
    assert len(ta) == len(tb)
    rel = absd = 0.0
    for x, y in zip(ta, tb):
        assert x.shape == y.shape and x.dtype == y.dtype
        if torch.equal(x, y):
            continue
        d = (x - y).abs().double()
        absd = max(absd, d.max().item())
        rel = max(rel, (d.norm() / max(y.abs().double().norm().item(), 1e-300)).item())
    return rel, absd


@pytest.mark.parametrize("material,start", [("indexed", "random"), ("dense", "random"), ("indexed", "rest"),
                                            ("dense", "rest")])
def test_single_pass_matches_the_oracle_and_the_two_kernel_step(gpu, material, start):
    """The single-pass step (lsf_set_fused) on a CPML box (kappa and alpha, dipole, DFT, odd stride) of 136 x 28 x 40
    cells, whose tiles fall in all three CPML classes (none, x only, other): random fields in a dielectric box, or
    D-form Lorentz nodes driven from rest by the dipole (the D-form update holds D = eps_inf E + sum P, which the
    reference's zero D and E_prev match only for fields at rest). Against the FP64 oracle to the parity tolerance;
    against the two-kernel step to rounding (after the run within 2e-5 relative L2 on every state tensor, the pole
    state being the most sensitive); eager stepping and graph replay (two graphs for the odd stride) bit-identical."""
    steps = 140
    kw = dict(shape=(136, 28, 40), n_steps_hint=steps, material=material, kappa_max=2.0, alpha=0.02, stride=7,
              frac=0.0 if start == "random" else 0.1)
    ns = cases.product_ns()
    a = cases.dispersive_box(ns, gpu, fused=True, **kw)
    b = cases.dispersive_box(ns, gpu, fused=True, **kw)
    ref = cases.dispersive_box(ns, gpu, fused=False, **kw)
    o = cases.dispersive_box(cases.oracle_ns(), **{k: v for k, v in kw.items() if k != "material"})
    assert a._two_sets and not ref._two_sets
    if start == "random":
        for e in (a, b, ref, o):
            cases.random_fields(e, 5)
    a.steps(1)
    ref.steps(1)
    torch.cuda.synchronize()
    rel1, abs1 = _state_diff(a, ref)
    a.steps(a.stride - 1)
    ref.steps(ref.stride - 1)
    o.chunk()
    a.sample()
    ref.sample()
    for _ in range(steps // a.stride - 1):
        a.chunk()
        ref.chunk()
        o.chunk()
    torch.cuda.synchronize()
    assert a.parity == (steps % 2)
    run, used = b.make_runner(True)
    assert used
    for _ in range(steps // b.stride):
        run()
    b.release()
    torch.cuda.synchronize()
    same = all(torch.equal(x, y) for x, y in zip(a.state_tensors(), b.state_tensors()))
    rel, absd = _state_diff(a, ref)
    print(f"single pass ({material}, {start}) vs two kernels: after 1 step rel {rel1:.1e} max {abs1:.1e}; after "
          f"{steps} steps rel {rel:.1e} max {absd:.1e}; graphs bit-identical to eager: {same}")
    _compare(o, a, f"single pass ({material}, {start}) vs FP64 oracle, {steps} steps")
    eo, ho = o.monitors[0].results()
    ea, ha = a.monitors[0].results()
    assert cases.rel_l2(ea, eo) < DFT_TOL and cases.rel_l2(ha, ho) < DFT_TOL
    assert same
    assert rel1 < 1e-7 and rel < 2e-5


def test_ade_media_table_is_validated(gpu):
    """lsf_set_ade refuses a D-form media table whose pole count is not an integer in [0, max_poles] or whose
    coefficients past a medium's poles are not zero (the E kernel advances max_poles poles for every node), and
    accepts the engine's own table again afterwards."""
    from lightsmith_amd import _lib
    eng = cases.dispersive_box(cases.product_ns(), gpu, shape=(20, 18, 16), with_dipole=False, monitors=False)
    lib, mp = eng.lib, eng.ade_max_poles
    assert mp == 3 and eng.n_ade > 0

    def set_ade(tab):
        return lib.lsf_set_ade(eng._h, _lib.ptr(eng._ade_chunk), _lib.ptr(eng._ade_P), eng.n_ade, mp, _lib.ptr(tab),
                               len(eng.ade_media))

    bad = eng._ade_tab.clone()
    bad[0, 1] = 2.0  # two poles declared, the third pole's coefficients still set
    assert set_ade(bad) == _lib.LSF_EINVAL and b"past a medium's poles" in lib.lsf_last_error()
    bad = eng._ade_tab.clone()
    bad[0, 1] = 2.5
    assert set_ade(bad) == _lib.LSF_EINVAL and b"integer" in lib.lsf_last_error()
    assert set_ade(eng._ade_tab) == _lib.LSF_OK
    eng.close()


@pytest.mark.parametrize("pole_counts", [(1,), (2,), (4,), (1, 3), (4, 2, 1)])
def test_pole_counts_and_mixed_media(gpu, pole_counts):
    """The E kernel is instantiated per exact pole count (the engine's max_poles); media with fewer poles than that
    share it with zero coefficients in their unused poles. Lorentz media of 1, 2 and 4 poles alone, and mixtures in
    one engine (one pure group per medium), against the FP64 oracle at the parity tolerance."""
    steps = 200
    o = cases.dispersive_box(cases.oracle_ns(), n_steps_hint=steps, pole_counts=pole_counts)
    g = cases.dispersive_box(cases.product_ns(), gpu, n_steps_hint=steps, pole_counts=pole_counts)
    assert g.ade_max_poles == max(pole_counts) and len(g.ade_media) == len(pole_counts)
    for _ in range(steps // o.stride):
        o.chunk()
        g.chunk()
    _compare(o, g, f"pole counts {pole_counts}, {steps} steps")


@pytest.mark.parametrize("pole_counts", [(3,), (2, 3)])
def test_pole_pairs_complex_bloch(gpu, pole_counts):
    """Complex64 fields, Bloch-periodic in x and y, with Lorentz D-form nodes of 2 and 3 poles (a complex node's
    pole pairs move in one 16-byte access each): random E and H on the stepped cells, zero E at the D-form nodes
    (their pole state starts at rest), against the FP64 oracle."""
    shape, dx, steps = (22, 18, 30), 2.0, 120
    periodic, kpar = (True, True, False), (0.06, -0.04, 0.0)
    rng = np.random.default_rng(11)
    eps = rng.choice([1.0, 2.25, 4.0], size=shape)
    inv_eps = np.broadcast_to(1.0 / eps, (3,) + shape).copy()
    disp = np.zeros(shape, bool)
    disp[2:-2, 2:-2, 8:-8] = rng.random((shape[0] - 4, shape[1] - 4, shape[2] - 16)) < 0.15
    ii, jj, kk = np.nonzero(disp)
    inv_eps.reshape(3, -1)[:, np.ravel_multi_index((ii, jj, kk), shape)] = 1.0
    which = rng.integers(0, len(pole_counts), size=len(ii))
    kw = dict(shape=shape, dx=dx, dt=0.9 * dx / np.sqrt(3), inv_eps=inv_eps, periodic=periodic, bloch=kpar)
    engines = []
    for ns, dev in ((cases.oracle_ns(), None), (cases.product_ns(), gpu)):
        groups = []
        for q, npl in enumerate(pole_counts):
            sel = which == q
            m = cases.medium(ns, 1.0, cases.lorentz3(np.random.default_rng(40 + q), npl))
            gi = np.sort(np.concatenate([ns.flat_index(c, ii[sel], jj[sel], kk[sel], shape) for c in range(3)]))
            groups.append(ns.DispersiveGroup(idx=gi, media=[m], frac=np.ones((1, len(gi))), w=None))
        pml = ns.PMLSpec(cells=((0, 0), (0, 0), (6, 6)))
        engines.append(cases.make(ns, dev, pml=pml, groups=groups, **kw))
    o, g = engines
    assert g.ade_max_poles == max(pole_counts)
    frng = np.random.default_rng(5)
    for c in range(3):
        for F, R in (("E", o.RE[c]), ("H", o.RH[c])):
            size = tuple(r.stop - r.start for r in R)
            vals = frng.standard_normal(size) + 1j * frng.standard_normal(size)
            if F == "E":
                vals[disp[R]] = 0.0
            getattr(o, F)[c][R] = vals
            getattr(g, F)[c][R] = torch.as_tensor(vals, dtype=torch.complex64, device=g.E_flat.device)
    o.run(steps)
    g.steps(steps)
    _compare(o, g, f"complex Bloch, pole counts {pole_counts}, {steps} steps", mask=cases.real_mask(shape, periodic))


@pytest.mark.parametrize("tiling", [(1, 4, 64, 5), (2, 4, 64, 16), (1, 8, 32, 3)])
def test_tilings_agree_with_poles(gpu, tiling):
    """The E kernel of an engine with pole state marches twice the tiling's z planes (10, 32, 6 here, partial last
    tiles included) while the H kernel marches the tiling's own: 3-pole D-form nodes in a CPML box against the FP64
    oracle under tilings other than the default."""
    steps = 120
    o = cases.dispersive_box(cases.oracle_ns(), n_steps_hint=steps)
    ns = cases.product_ns()
    old = ns.Engine.tiling
    ns.Engine.tiling = tiling
    try:
        g = cases.dispersive_box(ns, gpu, n_steps_hint=steps)
    finally:
        ns.Engine.tiling = old
    assert g.ade_max_poles == 3
    for _ in range(steps // o.stride):
        o.chunk()
        g.chunk()
    _compare(o, g, f"tiling {tiling} with 3-pole nodes, {steps} steps")


def test_table_source_over_a_substrate(gpu):
    """The table-fed plane wave with a layered background (dielectric substrate crossing the box, a wave coming up
    from it at 40 degrees): the HIP engine fed with the tables against the FP64 oracle fed with the same ones, fields
    after 900 steps (the pulse peak reaches the box around step 700) and the incident-wave view's spectrum."""
    from lightsmith_amd.fdtd import layered
    from lightsmith_amd.fdtd.media import Medium
    from oracle import sources as osrc
    shape, dx, eps_h, eps_s = (30, 28, 34), 4.0, 1.69, 2.25
    dt = 0.95 * dx / np.sqrt(3) * np.sqrt(eps_h)
    z_int = 16.0 * dx
    z = np.arange(shape[2]) * dx
    f_t, f_n = np.clip((z_int - (z - dx / 2)) / dx, 0, 1), np.clip((z_int - z) / dx, 0, 1)
    inv = np.empty((3,) + shape)
    inv[0] = inv[1] = 1 / (f_t * eps_s + (1 - f_t) * eps_h)
    inv[2] = f_n / eps_s + (1 - f_n) / eps_h
    th = np.deg2rad(40.0)
    khat = np.array([np.sin(th), 0.0, np.cos(th)])
    om = 2 * np.pi / np.array([280.0, 400.0])
    src = layered.ObliqueLayeredPlaneWave(shape=shape, origin=np.zeros(3), spacing=dx, dt=dt, lo=(9, 8, 10),
                                          hi=(21, 20, 24), khat=khat, ehat=(0.0, 1.0, 0.0), inv_eps=inv,
                                          host=Medium(eps_h), z_int=z_int, sub=Medium(eps_s),
                                          omega_c=om.mean(), omega_half=0.3 * om.mean(), max_steps=2000,
                                          omega_min=om.min())
    so = osrc.TableSource(src.T.reshape(-1), src.m, src.cmax, src.points, inc=src.inc,
                          T_inc=np.asarray(src.T_inc).reshape(-1))
    on, pn = cases.oracle_ns(), cases.product_ns()
    mo = on.DFTMonitor(om, np.arange(3), source=so)
    mp = pn.DFTMonitor(om, np.arange(3), source=src)
    kw = dict(shape=shape, dx=dx, dt=dt, inv_eps=inv, dft_stride=5)
    o = cases.make(on, pml=on.PMLSpec(cells=((5, 5),) * 3, eps_ref=eps_s), sources=[so], monitors=[mo], **kw)
    g = cases.make(pn, gpu, pml=pn.PMLSpec(cells=((5, 5),) * 3, eps_ref=eps_s), sources=[src], monitors=[mp], **kw)
    for _ in range(900 // 5):
        o.chunk()
        g.chunk()
    _compare(o, g, "table-fed TFSF over a substrate")
    eo, _ = o.monitors[0].results()
    eg, _ = g.monitors[0].results()
    assert cases.rel_l2(eg, eo) < DFT_TOL
