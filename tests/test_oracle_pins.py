"""What pins the FP64 oracle (CPU): closed forms, invariants and brute force, each chosen so that a
plausible mistake (dropped term, wrong sign or index, transposed operand) fails it. The reference
ships no golden vectors for this path (its tests compare against Mie/TMM/closed forms), so these
checks stand in for them.
"""

import numpy as np
import pytest

from oracle import fdtd, media, sources

DT = 0.95 / np.sqrt(3)


def _energy(eng, eps, h_old):
    return eps * np.sum(eng.E.real ** 2) + np.sum(h_old.real * eng.H.real)


def test_pec_box_conserves_discrete_energy():
    """Lossless PEC box: W = eps E^n.E^n + H^{n-1/2}.H^{n+1/2} is conserved to rounding (the restatement of
    test_fdtd_graded.py:200-237 on a uniform grid): the two curls are adjoint, so a wrong index or sign
    in either curl breaks it."""
    shape, eps = (11, 13, 12), 2.25
    eng = fdtd.Engine(shape=shape, dx=1.0, dt=DT, inv_eps=np.full((3,) + shape, 1 / eps))
    rng = np.random.default_rng(1)
    for c in range(3):
        eng.E[c][eng.RE[c]] = rng.standard_normal(eng.E[c][eng.RE[c]].shape)
        eng.H[c][eng.RH[c]] = rng.standard_normal(eng.H[c][eng.RH[c]].shape)
    W = []
    for _ in range(1500):
        h_old = eng.H.copy()
        eng.update_H()
        W.append(_energy(eng, eps, h_old))
        eng.update_E()
    W = np.asarray(W)
    assert np.abs(W / W[0] - 1).max() < 1e-10


def _recurrence_residual(series, cos_wdt):
    s = np.asarray(series)
    r = s[2:] - 2 * cos_wdt * s[1:-1] + s[:-2]
    return np.abs(r).max() / np.abs(s).max()


def test_pec_cavity_mode_has_the_discrete_eigenfrequency():
    """TM110 mode E_z = sin(pi i / (n0-1)) sin(pi j / (n1-1)) of a PEC box is an exact eigenmode of the Yee
    update: E^{n+1} - 2 cos(w dt) E^n + E^{n-1} = 0 with sin^2(w dt/2) = dt^2/(4 eps) (Kx^2 + Ky^2),
    K = 2 sin(k d / 2) / d."""
    shape, eps, dx, dt = (17, 13, 6), 2.0, 1.5, 0.9 * 1.5 / np.sqrt(3)
    eng = fdtd.Engine(shape=shape, dx=dx, dt=dt, inv_eps=np.full((3,) + shape, 1 / eps))
    i, j = np.arange(shape[0]), np.arange(shape[1])
    mode = np.outer(np.sin(np.pi * i / (shape[0] - 1)), np.sin(np.pi * j / (shape[1] - 1)))
    eng.E[2][:, :, :-1] = mode[:, :, None]
    K2 = sum((2 * np.sin(np.pi / (2 * (n - 1))) / dx) ** 2 for n in shape[:2])
    s2 = dt ** 2 / (4 * eps) * K2
    cos_wdt = 1 - 2 * s2
    series = []
    for _ in range(300):
        series.append(eng.E[2][5, 4, 2])
        eng.step()
    assert _recurrence_residual(series, cos_wdt) < 1e-12
    assert abs(_recurrence_residual(series, cos_wdt * (1 + 1e-4))) > 1e-6  # sensitive to the frequency


def test_bloch_mode_has_the_discrete_eigenfrequency():
    """A Bloch plane wave E_y = exp(i k x) in a cell periodic along x (Bloch k), y and z: its ghost planes
    carry exp(+-i k L) (engine.py:595-611), so it is an exact eigenmode with
    sin^2(w dt/2) = dt^2/eps (sin(k dx/2)/dx)^2. A wrong ghost phase sign makes it no eigenmode."""
    shape, eps, dx = (12, 5, 5), 1.7, 1.0
    k = 0.37
    dt = 0.9 / np.sqrt(3)
    eng = fdtd.Engine(shape=shape, dx=dx, dt=dt, inv_eps=np.full((3,) + shape, 1 / eps),
                      periodic=(True, True, True), bloch=(k, 0.0, 0.0))
    x = np.arange(shape[0]) * dx
    eng.E[1][:] = np.exp(1j * k * x)[:, None, None]
    s2 = dt ** 2 / eps * (np.sin(k * dx / 2) / dx) ** 2
    series = []
    for _ in range(200):
        series.append(eng.E[1][4, 2, 2])
        eng.step()
    assert _recurrence_residual(series, 1 - 2 * s2) < 1e-12
    bad = fdtd.Engine(shape=shape, dx=dx, dt=dt, inv_eps=np.full((3,) + shape, 1 / eps),
                      periodic=(True, True, True), bloch=(-k, 0.0, 0.0))
    bad.E[1][:] = np.exp(1j * k * x)[:, None, None]
    series = []
    for _ in range(200):
        series.append(bad.E[1][4, 2, 2])
        bad.step()
    assert _recurrence_residual(series, 1 - 2 * s2) > 1e-3


class _Point:
    """A one-node stand-in engine for driving an ADE group directly."""

    def __init__(self, dt, n=1):
        self.dt = dt
        self.fdt = np.float64
        self.E_flat = np.zeros(n)


@pytest.mark.parametrize("poles", [
    [(0.8, 0.0, 0.05)],                               # Drude
    [(0.5, 0.6, 0.07), (0.3, 1.1, 0.2)],              # two Lorentz poles
    [(0.9, 0.0, 0.04), (0.4, 0.9, 0.1), (0.2, 1.6, 0.3)],
])
def test_ade_realises_the_discrete_permittivity(poles):
    """Driving a pure D-form group with D^n = Re(D0 Z^n), Z = exp(-i w dt), its E settles to Re(D0 Z^n / eps_d)
    with eps_d = eps_inf + sum c3 / (Z - c1 - c2 / Z) (layered.py:24-31): pins c1, c2, c3 and the order
    of the D, P and E updates (engine.py:108-111, 769-793)."""
    m = media.Medium(1.8, tuple(media.Pole(*p) for p in poles))
    dt, w = 0.5, 0.45
    g = fdtd.PureGroup(_Point(dt), fdtd.DispersiveGroup(idx=np.array([0]), media=[m], frac=np.ones((1, 1)), w=None))
    Z = np.exp(-1j * w * dt)
    D0 = 1.0 + 0.3j
    eng = g and _Point(dt)
    d_prev = 0.0
    out = []
    n_steps = 6000
    for n in range(1, n_steps + 1):
        d_now = (D0 * Z ** n).real
        eng.E_flat[0] = g.E_prev[0] + (d_now - d_prev)
        g.step(eng)
        d_prev = d_now
        out.append(eng.E_flat[0])
    eps_d = media.discrete_eps(m, w, dt)
    n = np.arange(1, n_steps + 1)
    expect = (D0 * Z ** n / eps_d).real
    tail = slice(n_steps - 400, n_steps)
    err = np.abs(np.asarray(out)[tail] - expect[tail]).max() / np.abs(expect[tail]).max()
    assert err < 1e-8
    # the continuous eps at the same frequency differs: the check resolves the discretisation
    assert abs(m.eps(w) - eps_d) > 1e-4


def test_amplification_is_bounded_at_the_chosen_dt():
    """time_step lowers dt until one step of the Yee + ADE update is non-amplifying (setup.py:26-91;
    test_fdtd.py:102-120)."""
    m = media.drude_lorentz(1.0, 9.0, 0.07, [(2.0, 3.0, 0.5)])
    for dx in (40.0, 5.0):
        dt = media.time_step(dx, 0.95, [m])
        assert media.amplification(m, dt, np.linspace(0, 12 / dx ** 2, 401)) <= 1 + 1e-6


def _dipole_run(shape, pml, pos, steps, probe):
    inv = np.ones((3,) + shape)
    src = sources.PointDipole(shape=shape, origin=np.zeros(3), dx=1.0, dt=DT, pos=pos, pdir=(0, 0, 1.0),
                              inv_eps=inv, omega_c=0.5, omega_half=0.25, max_steps=steps + 5)
    eng = fdtd.Engine(shape=shape, dx=1.0, dt=DT, inv_eps=inv,
                      pml=fdtd.PMLSpec(cells=((pml, pml),) * 3) if pml else None, sources=[src])
    out = []
    for _ in range(steps):
        eng.step()
        out.append(eng.E[2][probe])
    return np.asarray(out)


def test_cpml_reflection_is_small_against_a_large_domain():
    """Brute force: a pulse from a dipole seen by a probe next to a 6-cell CPML, against the same probe in a
    domain so large that its walls' echo has not arrived. The difference is the CPML's reflection."""
    small, big = (30, 30, 30), (74, 74, 74)
    steps = 150
    a = _dipole_run(small, 6, (15.0, 15.0, 15.2), steps, (8, 15, 15))
    b = _dipole_run(big, 0, (37.0, 37.0, 37.2), steps, (30, 37, 37))
    rel = np.abs(a - b).max() / np.abs(b).max()
    assert rel < 5e-3
    # without the CPML the PEC wall reflects strongly in the same window
    c = _dipole_run(small, 0, (15.0, 15.0, 15.2), steps, (8, 15, 15))
    assert np.abs(c - b).max() / np.abs(b).max() > 0.1


def test_dft_sample_phase_convention():
    """acc_e += exp(+i w t) E(t) at t = n dt; acc_h uses t - dt/2 (engine.py:564-592). A sine over whole periods
    sampled this way gives +i N/2 exactly when the sign and the H half step are right."""
    N, stride = 64, 1
    dt = 0.25
    w0 = 2 * np.pi * 4 / (N * dt)
    shape = (3, 3, 3)
    eng = fdtd.Engine(shape=shape, dx=1.0, dt=dt, inv_eps=np.ones((3,) + shape))
    mon = fdtd.DFTMonitor(np.array([w0, 2 * w0]), np.array([13]), np.array([40]), np.array([41]))
    eng.monitors = [mon]
    mon.acc_e = np.zeros((2, 1), complex)
    mon.acc_h = np.zeros((2, 1), complex)
    for n in range(1, N + 1):
        eng.n = n
        eng.E_flat[13] = np.sin(w0 * n * dt)
        eng.H_flat[40] = eng.H_flat[41] = np.sin(w0 * (n - 0.5) * dt)
        eng.sample()
    e, h = mon.results()
    assert abs(e[0, 0] - 0.5j * N) < 1e-9 and abs(h[0, 0] - 0.5j * N) < 1e-9
    assert abs(e[1, 0]) < 1e-9


def test_dipole_current_has_no_net_charge_and_unit_weights():
    """j is the exact difference of g over omega_c dt, so sum j dt = (g(end) - g(0)) / omega_c ~ 0
    (emission.py:120-127); the trilinear weights of each component sum to its direction cosine."""
    shape = (20, 20, 20)
    src = sources.PointDipole(shape=shape, origin=np.zeros(3), dx=2.0, dt=DT * 2, pos=(20.3, 19.1, 21.7),
                              pdir=(0.6, 0.0, 0.8), inv_eps=np.ones((3,) + shape), omega_c=0.3,
                              omega_half=0.05, max_steps=3000)
    assert abs(np.sum(src.wave)) < 1e-9 * np.abs(src.wave).max() * len(src.wave)
    comp = src.idx // int(np.prod(shape))
    assert abs(src.w[comp == 0].sum() - 0.6) < 1e-12 and abs(src.w[comp == 2].sum() - 0.8) < 1e-12
    assert not np.any(comp == 1)


def test_tfsf_box_does_not_leak_axis_aligned():
    """Axis-aligned incidence: the 1D line's nodes coincide with the 3D Yee positions, so the scattered-field
    region of an empty box stays at rounding level (test_fdtd.py:179-200 is this test at solve level)."""
    shape = (26, 26, 30)
    inv = np.full((3,) + shape, 1 / 1.77)
    src = sources.TFSFPlaneWave(shape=shape, origin=np.zeros(3), dx=1.0, dt=DT, r0=np.array([13.0, 13, 15]),
                                lo=np.array([8, 8, 8]), hi=np.array([18, 18, 22]), khat=(0, 0, 1.0),
                                ehat=(1.0, 0, 0), eps_b=1.77, omega_c=0.6, omega_half=0.2, max_steps=400)
    eng = fdtd.Engine(shape=shape, dx=1.0, dt=DT, inv_eps=inv, pml=fdtd.PMLSpec(cells=((5, 5),) * 3, eps_ref=1.77),
                      sources=[src])
    peak_tot, peak_sca = 0.0, 0.0
    inside = (slice(None), slice(10, 17), slice(10, 17), slice(10, 21))
    out = np.ones(shape, bool)
    out[7:20, 7:20, 7:24] = False
    for _ in range(250):
        eng.step()
        peak_tot = max(peak_tot, np.abs(eng.E[inside]).max())
        peak_sca = max(peak_sca, np.abs(eng.E[:, out]).max())
    assert peak_tot > 0.1
    assert peak_sca < 1e-12 * peak_tot


def test_tfsf_oblique_leak_is_small():
    """Oblique incidence through the interpolated 1D line leaks ~1e-3 of the incident field (tfsf.py:8-10)."""
    shape = (26, 26, 26)
    inv = np.ones((3,) + shape)
    k = np.array([1.0, 2.0, 3.0]) / np.sqrt(14)
    e = np.cross(k, [0, 1.0, 0])
    e /= np.linalg.norm(e)
    src = sources.TFSFPlaneWave(shape=shape, origin=np.zeros(3), dx=1.0, dt=DT, r0=np.array([13.0, 13, 13]),
                                lo=np.array([8, 8, 8]), hi=np.array([18, 18, 18]), khat=k, ehat=e, eps_b=1.0,
                                omega_c=0.5, omega_half=0.2, max_steps=400)
    eng = fdtd.Engine(shape=shape, dx=1.0, dt=DT, inv_eps=inv, pml=fdtd.PMLSpec(cells=((5, 5),) * 3), sources=[src])
    out = np.ones(shape, bool)
    out[6:21, 6:21, 6:21] = False
    peak_tot = peak_sca = 0.0
    for _ in range(250):
        eng.step()
        peak_tot = max(peak_tot, np.abs(eng.E[:, 10:17, 10:17, 10:17]).max())
        peak_sca = max(peak_sca, np.abs(eng.E[:, out]).max())
    assert peak_tot > 0.1
    assert 1e-8 < peak_sca / peak_tot < 2e-2


def test_bloch_plane_source_does_not_leak():
    """The auxiliary Bloch column satisfies the main grid's discrete equations exactly (tfsf.py:214-222), so
    the scattered-field side of the plane stays at rounding level at any k_par and polarisation."""
    nz = 60
    shape = (8, 7, nz)
    sp = (20.0, 24.0, 15.0)
    kpar = (0.004, 0.002)
    dt = 0.9 / np.sqrt(sum(1 / d ** 2 for d in sp))
    pml = fdtd.PMLSpec(cells=((0, 0), (0, 0), (8, 8)), eps_ref=1.0)
    for pol in ("s", "p"):
        src = sources.BlochPlaneSource(shape=shape, spacing=sp, dt=dt, k_src=25, k_launch=24, kpar=kpar, pol=pol,
                                       eps_b=1.0, pml=pml, omega_c=2 * np.pi / 600, omega_half=0.3 * 2 * np.pi / 600,
                                       max_steps=900)
        eng = fdtd.Engine(shape=shape, dx=sp, dt=dt, inv_eps=np.ones((3,) + shape), pml=pml, sources=[src],
                          periodic=(True, True, False), bloch=(*kpar, 0.0))
        peak_tot = peak_sca = 0.0
        for _ in range(500):
            eng.step()
            peak_tot = max(peak_tot, np.abs(eng.E[:, 1:-1, 1:-1, 27:50]).max())
            peak_sca = max(peak_sca, np.abs(eng.E[:, 1:-1, 1:-1, 9:23]).max())
        assert peak_tot > 1e-3
        assert peak_sca < 1e-11 * peak_tot, pol


def test_mixed_group_reduces_to_pure():
    """One dispersive medium at fraction 1: both paths of the mixed update are the pure update, whatever w
    (engine.py:722-728)."""
    m = media.Medium(2.0, (media.Pole(0.5, 0.4, 0.05), media.Pole(0.2, 1.3, 0.1)))
    n = 5
    w = np.linspace(0, 1, n)
    rng = np.random.default_rng(0)
    drive = rng.standard_normal((200, n))
    pure = fdtd.PureGroup(_Point(0.5, n), fdtd.DispersiveGroup(idx=np.arange(n), media=[m], frac=np.ones((1, n)), w=None))
    mixed = fdtd.MixedGroup(_Point(0.5, n), fdtd.DispersiveGroup(idx=np.arange(n), media=[m], frac=np.ones((1, n)), w=w))
    ea, eb = _Point(0.5, n), _Point(0.5, n)
    for row in drive:
        ea.E_flat[:] = pure.E_prev + row
        eb.E_flat[:] = mixed.E_prev + row
        pure.step(ea)
        mixed.step(eb)
    assert np.allclose(ea.E_flat, eb.E_flat, rtol=1e-12, atol=1e-14)


def test_quintic_table_read_is_exact_for_quintic_polynomials():
    """read_tables interpolates with the 6-point Lagrange stencil -2..3 (layered.py:164-187): samples of a degree-5
    polynomial are reproduced exactly at every fractional position, and the integer/fraction split of the
    position (n + i0 + u0) / m = c + u is the one the stencil weights assume."""
    m, L = 3, 60
    coef = np.array([0.3, -1.2, 0.7, 0.05, -0.011, 0.0004])
    poly = lambda x: sum(a * x ** p for p, a in enumerate(coef))  # noqa: E731
    T = np.concatenate([poly(np.arange(L, dtype=float)), np.zeros(8)])
    rng = np.random.default_rng(3)
    i0 = rng.integers(6, 60, 200)
    u0 = rng.uniform(0, 1, 200)
    base = np.full(200, -2)
    for n in (0, 5, 17, 40):
        got = sources.read_tables(T, n, m, L - 6, base, i0, u0)
        x = (n + i0 + u0) / m
        ok = x < L - 6
        assert np.allclose(got[ok], poly(x[ok]), rtol=1e-12, atol=1e-9)
    # a degree-6 term is not reproduced: the check sees the stencil order
    T6 = np.concatenate([np.arange(L, dtype=float) ** 6, np.zeros(8)])
    got = sources.read_tables(T6, 0, m, L - 6, base, i0, u0)
    x = (i0 + u0) / m
    assert np.abs(got - x ** 6).max() > 1e-3


def test_mixed_group_static_limit_is_the_kottke_average():
    """Non-dispersive media only: the mixed update maps D to E with w <1/eps> + (1 - w) / <eps> (the anisotropic
    subpixel average, engine.py:722-728)."""
    media_ = [media.Medium(1.5), media.Medium(4.0), media.Medium(11.0)]
    n = 6
    rng = np.random.default_rng(5)
    f = rng.dirichlet(np.ones(3), n).T
    w = rng.uniform(0, 1, n)
    g = fdtd.MixedGroup(_Point(0.5, n), fdtd.DispersiveGroup(idx=np.arange(n), media=media_, frac=f, w=w))
    eng = _Point(0.5, n)
    D = rng.standard_normal(n)
    eng.E_flat[:] = D
    g.step(eng)
    eps = np.array([m.eps_inf for m in media_])[:, None]
    expect = D * (w * np.sum(f / eps, axis=0) + (1 - w) / np.sum(f * eps, axis=0))
    assert np.allclose(eng.E_flat, expect, rtol=1e-13)


def test_mode_injection_is_the_modal_matrix_vector_product():
    shape = (6, 6, 6)
    eng = fdtd.Engine(shape=shape, dx=1.0, dt=0.3, inv_eps=np.ones((3,) + shape))
    rng = np.random.default_rng(0)
    h_idx = rng.choice(3 * 216, 10, replace=False)
    mH, tE = rng.standard_normal((10, 2)), rng.standard_normal((5, 2))
    src = sources.ModeInjection(h_idx, mH, tE, np.zeros(0, int), np.zeros((0, 2)), tE)
    src.bind(eng)
    eng.n = 3
    src.correct_H(eng)
    assert np.allclose(eng.H_flat[h_idx], mH @ tE[3])


def test_offdiagonal_coupling_is_symmetric():
    """The off-diagonal terms couple E_c at p and E_d at q with K(p, q) = (o_cd(p) + o_dc(q)) / 8 in both directions
    (engine.py:421-431): the operator recovered from the gather rows is symmetric, with each K entering the four
    curl taps of its source node."""
    shape = (9, 8, 7)
    rng = np.random.default_rng(8)
    oi, oj, ok = np.meshgrid(np.arange(2, 7), np.arange(2, 6), np.arange(2, 5), indexing="ij")
    idx3 = np.stack([oi.ravel(), oj.ravel(), ok.ravel()], axis=1)
    idx3 = np.concatenate([idx3] * 3)
    comp = np.repeat(np.arange(3), len(oi.ravel()))
    coef = rng.uniform(-0.1, 0.1, (len(comp), 3))
    eng = fdtd.Engine(shape=shape, dx=1.0, dt=0.5, inv_eps=np.ones((3,) + shape),
                      offdiag=fdtd.OffDiagonal(idx3=idx3, comp=comp, coef=coef))
    targets, h, w = eng.od
    N = int(np.prod(shape))
    K = {}
    for r, t in enumerate(targets):
        for slot in range(h.shape[1] // 4):
            if w[r, 4 * slot] == 0:
                continue
            cmp, rest = divmod(int(h[r, 4 * slot]), N)
            d = (cmp + 1) % 3  # the first tap reads H_{d+2} at the source node q
            s = d * N + rest
            K[(int(t), s)] = w[r, 4 * slot] / eng.dt
            # the four taps carry +K, -K, -K, +K
            assert np.allclose(w[r, 4 * slot:4 * slot + 4], np.array([1, -1, -1, 1]) * w[r, 4 * slot])
    assert len(K) > 100
    for (t, s), v in K.items():
        assert (s, t) in K and abs(K[(s, t)] - v) < 1e-15


# --- conformal vertex tensor (engine.py:839-923; the reference's own test: test_fdtd.py:254-284) --------------

def test_vertex_tensor_map_is_symmetric_and_bounded():
    """The D -> E map of the vertex tensor is a sum of congruences: symmetric, every Rayleigh quotient between
    1/eps_inf,max and 1/eps_min (the reference's test_fdtd_metal_vertex_tensor_is_symmetric_and_bounded)."""
    from oracle import media as omed
    from oracle.tensor import TensorGroup, vertex_tensor
    media = [omed.Medium(1.77, ()), omed.Medium(9.5, (omed.Pole(0.00208, 0.0, 1.8e-4),))]
    shape, c, r = (24, 24, 24), np.array([0.3, 0.1, -0.2]), 6.4
    tm = vertex_tensor(lambda p: (np.linalg.norm(p - c, axis=-1) < r).astype(int), media, shape, np.full(3, -12.0),
                       1.0, np.full(3, 2), np.full(3, 22))
    assert tm is not None and len(tm.normal) > 100
    eng = fdtd.Engine(shape=shape, dx=1.0, dt=0.5, inv_eps=np.ones((3,) + shape), groups=[], pml=None, sources=[],
                     monitors=[], dft_stride=1)
    idx = tm.idx

    def apply(x):
        g = TensorGroup(eng, tm)
        eng.E_flat[idx] = x
        g.step(eng)
        return np.array(eng.E_flat[idx])

    rng = np.random.default_rng(0)
    x, y = rng.standard_normal((2, len(idx)))
    mx, my = apply(x), apply(y)
    assert abs(x @ my - y @ mx) < 1e-12 * abs(x @ mx)
    for z in rng.standard_normal((5, len(idx))):
        q = z @ apply(z) / (z @ z)
        assert 1 / 9.5 <= q <= 1 / 1.77
    # planted mistake: taking the slot's own D for sigma (dropping the pair mean) breaks the symmetry
    g = TensorGroup(eng, tm)
    g.t_vc = g.t_vc * 0
    eng.E_flat[idx] = x
    g.step(eng)
    bx = np.array(eng.E_flat[idx])
    g = TensorGroup(eng, tm)
    g.t_vc = g.t_vc * 0
    eng.E_flat[idx] = y
    g.step(eng)
    by = np.array(eng.E_flat[idx])
    assert abs(x @ by - y @ bx) > 1e-6 * abs(x @ bx)


def test_vertex_tensor_in_one_medium_is_the_plain_update():
    """With every vertex and plain slot in one non-dispersive medium the map is D / eps for every node."""
    from oracle import media as omed
    from oracle.tensor import TensorGroup, vertex_tensor
    media = [omed.Medium(1.77, ()), omed.Medium(9.5, (omed.Pole(0.00208, 0.0, 1.8e-4),))]
    shape, c, r = (20, 20, 20), np.zeros(3), 5.2
    tm = vertex_tensor(lambda p: (np.linalg.norm(p - c, axis=-1) < r).astype(int), media, shape, np.full(3, -10.0),
                       1.0, np.full(3, 2), np.full(3, 18))
    tm.frac[:] = 0.0
    tm.frac[0] = 1.0
    tm.slot_medium[:] = 0
    eng = fdtd.Engine(shape=shape, dx=1.0, dt=0.5, inv_eps=np.ones((3,) + shape), groups=[], pml=None, sources=[],
                     monitors=[], dft_stride=1)
    g = TensorGroup(eng, tm)
    x = np.random.default_rng(1).standard_normal(len(tm.idx))
    eng.E_flat[tm.idx] = x
    g.step(eng)
    assert np.max(np.abs(eng.E_flat[tm.idx] - x / 1.77)) < 1e-14


def test_mie_series_reproduces_the_worked_example_and_limits():
    """oracle/mie.py: Bohren & Huffman's BHMIE example (m = 1.55, x = 5.213: Q_ext = Q_sca = 3.10543,
    Q_back = 2.92534), the Rayleigh limit Q_sca = 8/3 x^4 |(m^2 - 1)/(m^2 + 2)|^2 and the optical theorem
    (Q_ext = Q_sca without loss, Q_abs > 0 with it)."""
    from oracle.mie import cross_sections, mie_q
    qe, qs, qb = mie_q(1.55, 2 * np.pi * 0.525 / 0.6328)
    assert abs(qe - 3.10543) < 2e-5 and abs(qs - 3.10543) < 2e-5 and abs(qb - 2.92534) < 2e-5
    for m in (1.5, 2.0 + 0.3j):
        x = 1e-3
        qe, qs, _ = mie_q(m, x)
        ray = 8 / 3 * x ** 4 * abs((m * m - 1) / (m * m + 2)) ** 2
        assert abs(qs / ray - 1) < 1e-5
    for x in (0.3, 2.0, 7.5):
        qe, qs, _ = mie_q(1.8, x)
        assert abs(qe - qs) < 1e-12 * max(1, qe)
        qe, qs, _ = mie_q(1.8 + 0.2j, x)
        assert qe - qs > 0
    ce, cs, ca = cross_sections(4.0, 60.0, [500.0], n_host=1.0)
    qe, qs, _ = mie_q(2.0, 2 * np.pi * 60.0 / 500.0)
    assert abs(ce[0] - qe * np.pi * 3600) < 1e-9 * ce[0] and abs(ca[0]) < 1e-9 * ce[0]


def test_tmm_matches_the_airy_sum_quarter_wave_mirror_and_conserves_energy():
    """oracle/tmm.py: a single film against the three-layer Airy sum (tests/test_drivers.py: a different method,
    itself pinned by Fresnel and the quarter-/half-wave cases) at normal and oblique incidence, s and p, lossless
    and Drude films; the quarter-wave mirror (HL)^N R = ((n0 - Y) / (n0 + Y))^2, Y = (n_H / n_L)^{2N} n_s;
    R + T = 1 for a lossless multilayer at every angle; the layer order matters (a swapped pair fails the mirror)."""
    from oracle.tmm import stack_rt
    from tests.test_drivers import thin_film
    wl = np.array([450.0, 600.0, 750.0])
    for n2 in (2.0, 0.3 + 3.1j):
        for th in (0.0, 0.52, 1.05):
            for pol in "sp":
                R, T = stack_rt([n2 ** 2], [37.0], 1.0, 2.25, wl, th, pol)
                for q, lam in enumerate(wl):
                    Ra, Ta = thin_film(1.0, n2, 1.5, 37.0, lam, th, pol)
                    assert abs(R[q] - Ra) < 1e-12 and abs(T[q] - Ta) < 1e-12
    nh, nl, ns, lam = 2.3, 1.38, 1.52, 550.0
    for N in (1, 3, 6):
        eps = [nh ** 2, nl ** 2] * N
        d = [lam / (4 * nh), lam / (4 * nl)] * N
        R, T = stack_rt(eps, d, 1.0, ns ** 2, [lam], 0.0, "s")
        Y = (nh / nl) ** (2 * N) * ns
        assert abs(R[0] - ((1 - Y) / (1 + Y)) ** 2) < 1e-12 and abs(R[0] + T[0] - 1) < 1e-12
        Rw, _ = stack_rt(eps[::-1], d[::-1], 1.0, ns ** 2, [lam], 0.0, "s")
        assert abs(Rw[0] - R[0]) > 1e-3
    rng = np.random.default_rng(3)
    eps = list(rng.uniform(1.2, 6.0, 5))
    d = list(rng.uniform(20, 200, 5))
    for th in (0.0, 0.4, 1.2):
        for pol in "sp":
            R, T = stack_rt(eps, d, 1.44, 2.1, np.linspace(400, 900, 7), th, pol)
            assert np.all(np.abs(R + T - 1) < 1e-12)


def test_rcwa_reduces_to_tmm_conserves_energy_and_converges():
    """oracle/rcwa.py: a uniform layer (and a uniform two-layer stack, lossy included) is the characteristic-matrix
    TMM for TE and TM at normal and oblique incidence, every other order empty; lossless gratings (a centred lamellar
    one and an asymmetric two-box one) conserve energy; N = 25 and N = 50 harmonics agree within 2e-3 on the
    lamellar grating of the reference's tests; a centred grating at normal incidence sends equal power into +n and
    -n; specular reflection of the asymmetric grating is reciprocal."""
    from oracle.rcwa import lamellar
    from oracle.tmm import stack_rt
    for pol in "sp":
        for th in (0.0, 0.3, 0.9):
            n, R, T = lamellar([(300.0, 4.0, [])], 1.0, 2.25, 700.0, 600.0, th, pol, 8)
            Rt, Tt = stack_rt([4.0], [300.0], 1.0, 2.25, [600.0], th, pol)
            assert abs(R[n == 0][0] - Rt[0]) < 1e-12 and abs(T[n == 0][0] - Tt[0]) < 1e-12
            assert R[n != 0].max() < 1e-20 and T[n != 0].max() < 1e-20
            n, R, T = lamellar([(40.0, -8.0 + 1.2j, []), (120.0, 2.1, [])], 1.44, 2.25, 500.0, 650.0, th, pol, 6)
            Rt, Tt = stack_rt([-8.0 + 1.2j, 2.1], [40.0, 120.0], 1.44, 2.25, [650.0], th, pol)
            assert abs(R[n == 0][0] - Rt[0]) < 1e-12 and abs(T[n == 0][0] - Tt[0]) < 1e-12
    sym = [(250.0, 1.0, [(0.0, 300.0, 6.0)])]
    asym = [(180.0, 1.5, [(-120.0, 200.0, 6.0), (190.0, 70.0, 3.0)])]
    for pol in "sp":
        for th in (0.0, 0.4):
            for lay in (sym, asym):
                n, R, T = lamellar(lay, 1.0, 2.25, 700.0, 600.0, th, pol, 20)
                assert abs(R.sum() + T.sum() - 1) < 1e-10
        a = lamellar(sym, 1.0, 2.25, 700.0, 600.0, 0.0, pol, 25)
        b = lamellar(sym, 1.0, 2.25, 700.0, 600.0, 0.0, pol, 50)
        for q in (1, 2):
            assert np.abs(a[q] - b[q][25:76]).max() < 2e-3
        n, R, T = a
        assert np.abs(R - R[::-1]).max() < 1e-12 and np.abs(T - T[::-1]).max() < 1e-12
        assert T[n == 1][0] > 0.1  # the first orders carry power
        Rp = lamellar(asym, 1.0, 2.25, 700.0, 600.0, 0.35, pol, 20)[1]
        Rm = lamellar(asym, 1.0, 2.25, 700.0, 600.0, -0.35, pol, 20)[1]
        assert abs(Rp[20] - Rm[20]) < 1e-9
        assert np.abs(Rp - Rp[::-1]).max() > 1e-3  # the asymmetric grating is not trivially symmetric


def test_rcwa_box_coefficients_are_the_profile_fft_and_the_subwavelength_limit_is_the_effective_medium():
    """The box Fourier coefficients equal the FFT of the profile sampled on 2^16 points (sign and centre of the
    phase, width in the sinc); a 30 nm period grating at 1000 nm is its effective film: TE sees <eps>, TM (E across
    the lines) 1 / <1/eps> -- the pair swapped misses by far more than the tolerance."""
    from oracle.rcwa import box_coefficients, lamellar
    from oracle.tmm import stack_rt
    L, nx = 700.0, 1 << 16
    boxes = [(-100.0, 180.0, 6.0), (220.0, 60.0, 3.0)]
    x = -L / 2 + (np.arange(nx) + 0.5) * L / nx
    prof = np.full(nx, 1.5)
    for c, w, e in boxes:
        prof[np.abs(x - c) < w / 2] = e
    m = np.arange(-10, 11)
    fft = (prof[None, :] * np.exp(-2j * np.pi * m[:, None] * x[None, :] / L)).mean(axis=1)
    assert np.abs(box_coefficients(1.5, boxes, L, m) - fft).max() < 2e-4
    inv = (1 / prof[None, :] * np.exp(-2j * np.pi * m[:, None] * x[None, :] / L)).mean(axis=1)
    assert np.abs(box_coefficients(1.5, boxes, L, m, inverse=True) - inv).max() < 2e-4
    f, e1, e2 = 0.4, 6.0, 1.0
    lay = [(200.0, e2, [(0.0, f * 30.0, e1)])]
    e_te, e_tm = f * e1 + (1 - f) * e2, 1 / (f / e1 + (1 - f) / e2)
    for pol, e_eff, e_wrong in (("s", e_te, e_tm), ("p", e_tm, e_te)):
        n, R, T = lamellar(lay, 1.0, 2.25, 30.0, 1000.0, 0.0, pol, 30)
        Re = stack_rt([e_eff], [200.0], 1.0, 2.25, [1000.0], 0.0, pol)[0][0]
        Rw = stack_rt([e_wrong], [200.0], 1.0, 2.25, [1000.0], 0.0, pol)[0][0]
        assert abs(R[n == 0][0] - Re) < 2e-3 and abs(Rw - Re) > 0.05


def test_dipole_substrate_model_reduces_to_mie_the_image_dipole_and_fresnel():
    """oracle/dipole_substrate.py: with the substrate equal to the host the sphere's C_sca is the Mie series' a_1
    term (both lit from above and from below); over a perfect conductor (r_s = -1, r_p = 1) the Sommerfeld integrals
    equal the image dipole's field at distance 2h in closed form (transverse image -p, axial image +p); the Fresnel
    amplitude driving the dipole reflects |r|^2 = TMM's R for s and p; a plausible slip (r_p with the eps ratio
    inverted) fails the last check."""
    from oracle.dipole_substrate import background_field, fresnel, reflected_green, sphere_above_substrate
    from oracle.mie import mie_ab
    from oracle.tmm import stack_rt
    wl = np.array([400.0, 600.0])
    for below in (False, True):
        c = sphere_above_substrate(20.0, 2.0, 1.0, 1.0, 40.0, wl, 30.0, "p", from_below=below)
        k = 2 * np.pi / wl
        a1 = np.array([mie_ab(2.0, kk * 20.0, 1)[0][0] for kk in k])
        assert np.allclose(c, 6 * np.pi / k ** 2 * np.abs(a1) ** 2, rtol=1e-9)
    k1, h = 2 * np.pi / 500.0, 60.0
    gp, gz = reflected_green(k1, h, lambda s: -1.0 + 0j, lambda s: 1.0 + 0j)
    d = 2 * h
    ph = np.exp(1j * k1 * d) / (4 * np.pi * d)
    assert abs(gp - (-ph * (k1 ** 2 + 1j * k1 / d - 1 / d ** 2))) < 1e-7 * abs(gp)
    assert abs(gz - 2 * ph * (1 / d ** 2 - 1j * k1 / d)) < 1e-7 * abs(gz)
    for th in (0.0, 0.5, 1.0):
        kx = np.sin(th)
        rs, rp = fresnel(1.0, 2.25, kx)
        for pol, r in (("s", rs), ("p", rp)):
            R = stack_rt([], [], 1.0, 2.25, [600.0], th, pol)[0][0]
            assert abs(abs(r) ** 2 - R) < 1e-12
        e, _ = background_field(1.0, 1.5, 2 * np.pi / 600.0, 2 * np.pi / 600.0 * kx, 0.0, "p", False)
        q1, q2 = np.cos(th), np.sqrt(2.25 - kx ** 2)
        wrong = (q1 - 2.25 * q2) / (q1 + 2.25 * q2)
        if th > 0:
            assert abs(abs(wrong) ** 2 - stack_rt([], [], 1.0, 2.25, [600.0], th, "p")[0][0]) > 1e-3
        assert abs(e[0]) > 0 and abs(e[1]) == 0  # p: E in the plane of incidence


def test_dipole_sphere_emission_obeys_the_optical_theorem_and_the_quasistatic_limit():
    """oracle/dipole_sphere.py: next to a lossless sphere the total rate equals the radiative rate (to 1e-12, for
    the radial and tangential forms and a mix); far from the sphere both are 1; a small lossy sphere quenches as
    1 + c R^3 Im(beta) / (k^3 d^6), c = 6 radial and 3/2 tangential, the error falling 4x per doubling of d (the
    test_emission.py:36-51 limits of the reference's own Mie emission)."""
    from oracle.dipole_sphere import dipole_emission
    for cos2 in (0.0, 0.3, 1.0):
        t, r = dipole_emission(60.0, 2.25 + 0j, [450.0, 600.0, 800.0], 1.0, 75.0, cos2)
        assert np.allclose(t, r, rtol=1e-12, atol=0) and np.all(np.abs(t - 1) > 0.05)
        t, r = dipole_emission(40.0, 4.0 + 0j, [500.0, 700.0], 1.33, 50.0, cos2)
        assert np.allclose(t, r, rtol=1e-12, atol=0)
    t, r = dipole_emission(60.0, 2.25 + 0j, [600.0], 1.0, 1e5, 1.0)
    assert abs(t[0] - 1) < 1e-4 and abs(r[0] - 1) < 1e-4
    eps, R, wl = 4 + 2j, 5.0, 1e5
    k, beta = 2 * np.pi / wl, (eps - 1) / (eps + 2)
    for cos2, c in ((1.0, 6.0), (0.0, 1.5)):
        err = []
        for d in (50.0, 100.0):
            tot, _ = dipole_emission(R, eps, [wl], 1.0, d, cos2)
            err.append((tot[0] - 1) / (c * R ** 3 * beta.imag / (k ** 3 * d ** 6)) - 1)
        assert abs(err[1]) < 0.01 and abs(err[0] / err[1] - 4) < 0.4


def test_mie_amplitudes_integrate_to_c_sca_and_obey_the_optical_theorem():
    """oracle/mie.py amplitudes: the angular functions' recurrence against their closed forms for n = 1, 2
    (pi_2 = 3 mu, tau_2 = 3 cos 2 theta); dC_sca/dOmega integrated over the sphere equals the series' C_sca (to the
    quadrature's 1e-6); C_ext = 4 pi / k^2 Re S(0) (the optical theorem) equals the series' C_ext; S1 = S2 forward
    and backward; a Rayleigh sphere's pattern is (cos^2 theta cos^2 phi + sin^2 phi) times |alpha|^2 k^4 / 16 pi^2."""
    from oracle.mie import amplitudes, cross_sections, dcsca_domega, mie_ab
    for eps, r, lam in ((2.25, 60.0, 500.0), (4.0 + 0.5j, 80.0, 600.0)):
        k = 2 * np.pi / lam
        th = np.linspace(0, np.pi, 2001)
        pat = dcsca_domega(eps, r, lam, np.degrees(th), [0.0, 90.0])
        c = 2 * np.pi * np.trapezoid(0.5 * (pat[:, 0] + pat[:, 1]) * np.sin(th), th)
        ce, cs, _ = cross_sections(eps, r, [lam])
        assert abs(c / cs[0] - 1) < 1e-5
        s1, s2 = amplitudes(np.sqrt(eps), k * r, [1.0, -1.0])
        assert abs(4 * np.pi / k ** 2 * s1[0].real / ce[0] - 1) < 1e-12
        assert abs(s1[0] - s2[0]) < 1e-12 * abs(s1[0]) and abs(s1[1] + s2[1]) < 1e-12 * abs(s1[1])
    x, m = 0.01, 1.5
    a, _ = mie_ab(m, x, 1)
    mu = np.cos(np.radians([0.0, 30.0, 90.0]))
    s1, s2 = amplitudes(m, x, mu)
    assert np.allclose(s1, 1.5 * a[0], rtol=1e-4) and np.allclose(s2, 1.5 * a[0] * mu, rtol=1e-4, atol=1e-11)  # b_1 ~ x^5


def test_mie_pole_finder_reproduces_the_quoted_poles_and_the_quasistatic_limit():
    """oracle/mie.py pole: the complex poles of a_1 and a_2 for a Drude sphere (eps_inf 9.5, 9 eV, 0.07 eV,
    r = 40 nm in water) at the values the reference quotes for them (test_fdtd.py:328-330: 533.8 nm, Q 8.74;
    487.5 nm, Q 35.1); a tiny sphere's a_1 pole is the quasi-static eps(omega) = -2 eps_host (5e-6)."""
    from oracle.mie import pole
    hbarc = 197.3269804
    wp, g = 9.0 / hbarc, 0.07 / hbarc

    def eps(w):
        return 9.5 - wp ** 2 / (w ** 2 + 1j * g * w)

    for n, lam, q, tol in ((1, 533.8, 8.74, 0.06), (2, 487.5, 35.1, 0.06)):
        target = -(n + 1) / n * 1.33 ** 2
        w0 = (-1j * g + np.sqrt(-g ** 2 + 4 * wp ** 2 / (9.5 - target))) / 2
        w = pole(n, eps, 40.0, 1.33, w0)
        assert abs(2 * np.pi / w.real - lam) < tol and abs(w.real / (-2 * w.imag) - q) < 0.01 * q
    w0 = (-1j * g + np.sqrt(-g ** 2 + 4 * wp ** 2 / 11.5)) / 2
    assert abs(pole(1, eps, 0.5, 1.0, w0 * 1.001) / w0 - 1) < 1e-5

