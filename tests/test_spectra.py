"""Layer stacks and periodic structures on the CPU (no solve): the problem types parse and dispatch, the backend
declines what the reference's FDTD declines (periodic.py:63-97), plans the unit cell by the reference's rules
(periodic.py:151-211) and counts its runs (periodic.py:266-302), the layer classifier puts every point in the
right layer with its periodic images (checked against brute force), the rasterised stack has the layered
averaging rules at its interfaces, and the order filter keeps the orders that propagate."""

import json
import math

import numpy as np
import pytest
import torch

from lightsmith_amd import api, backend
from lightsmith_amd import problem as P
from lightsmith_amd.fdtd import spectra

CI, CP = P.ConstantIndex, P.ConstantPermittivity
WL = P.WavelengthRange(start_nm=400, stop_nm=800, num=9)


def stack(material, thickness, pols=("s",), angles=(0.0,), **fdtd):
    return P.LayerStackProblem(superstrate=CI(n_real=1.0), substrate=CI(n_real=1.5),
                               layers=[P.Layer(thickness_nm=thickness, material=material)], wavelengths=WL,
                               polarizations=list(pols), angles_deg=list(angles), options={"fdtd": fdtd})


def lamellar(wl, pols=("s", "p"), theta=0.0, phi=0.0, **fdtd):
    return P.PeriodicProblem(
        lattice_a1_nm=(700, 0), superstrate=CI(n_real=1.0), substrate=CI(n_real=1.5),
        layers=[P.PatternedLayer(thickness_nm=250, background=CP(eps_real=1.0),
                                 shapes=[P.Structure2D(profile=P.Rectangle(size_nm=(300, 1.0)),
                                                       material=CP(eps_real=6.0))])],
        wavelengths=list(wl), incidence=P.Incidence(theta_deg=theta, phi_deg=phi, polarizations=list(pols)),
        options={"fdtd": fdtd})


def pillars(wl=(500.0, 700.0, 900.0), theta=0.0, phi=0.0, **fdtd):
    return P.PeriodicProblem(
        lattice_a1_nm=(400, 0), lattice_a2_nm=(0, 400), superstrate=CI(n_real=1.0), substrate=CI(n_real=1.45),
        layers=[P.PatternedLayer(thickness_nm=150, background=CI(n_real=1.0),
                                 shapes=[P.Structure2D(profile=P.Circle(radius_nm=120), material=CI(n_real=2.0))])],
        wavelengths=list(wl), incidence=P.Incidence(theta_deg=theta, phi_deg=phi), options={"fdtd": fdtd})


def test_problem_kinds_parse_and_dispatch():
    for prob in (stack(CI(n_real=2.0), 300), lamellar([600.0]), pillars()):
        back = P.parse_problem(json.loads(prob.model_dump_json()))
        assert type(back) is type(prob) and back == prob
    B = backend.NativeFDTD
    assert {"scattering", "periodic", "layer_stack"} <= set(B.problem_kinds)
    assert backend.get_backend("fdtd", kind="periodic") is B
    with pytest.raises(ValueError, match="kind"):
        P.parse_problem({"kind": "emission"})


def test_spectrum_difference_is_max_and_rms_per_quantity():
    kw = dict(method="fdtd", backend="x", wavelengths_nm=[500.0, 600.0], angles_deg=[0.0], polarizations=["s"])
    a = P.SpectrumResult(r={"s": [[0.1, 0.2]]}, t={"s": [[0.9, 0.8]]}, a={"s": [[0.0, 0.0]]}, **kw)
    b = P.SpectrumResult(r={"s": [[0.1, 0.5]]}, t={"s": [[0.9, 0.4]]}, a={"s": [[0.0, 0.1]]}, **kw)
    d = a.difference(b)
    assert d["r_s"]["max"] == pytest.approx(0.3) and d["t_s"]["max"] == pytest.approx(0.4)
    assert d["a_s"]["rms"] == pytest.approx(math.sqrt(0.01 / 2))


def test_supports_declines_like_the_reference():
    B = backend.NativeFDTD
    assert B.supports(lamellar([600.0]))[0]
    assert B.supports(lamellar([600.0], theta=60, phi=30))[0]
    ok, why = B.supports(lamellar([600.0], theta=88))
    assert not ok and "grazing" in why and "rcwa" in why
    ok, why = B.supports(stack(CI(n_real=2.0), 100).model_copy(update={"angles_deg": [0, 89]}))
    assert not ok and "tmm" in why
    assert not B.supports(lamellar([600.0]).model_copy(update={"lattice_a1_nm": (600, 100)}))[0]
    assert not B.supports(pillars().model_copy(update={"lattice_a2_nm": (50, 400)}))[0]
    assert B.supports(stack(CI(n_real=2.0), 100).model_copy(update={"substrate": CI(n_real=3.5, n_imag=0.1)}))[0]
    inc = stack(CI(n_real=2.0), 100).model_copy(
        update={"layers": [P.Layer(thickness_nm=100, material=CI(n_real=2.0), coherent=False)]})
    assert not B.supports(inc)[0]
    ok, why = B.supports(stack(CI(n_real=2.0), 100).model_copy(update={"superstrate": CI(n_real=1.0, n_imag=0.1)}))
    assert not ok and "superstrate" in why
    ok, why = B.supports(stack(CI(n_real=2.0), 100, precision="double"))
    assert not ok and "double" in why


def test_estimate_counts_runs_and_whole_cells_per_period():
    est = api.estimate(lamellar([600.0]))
    nx = est["real_cells"][0]
    assert est["spacing_nm"][0] * nx == pytest.approx(700.0)
    assert est["real_cells"][1] == 1 and est["runs"] == 2 and "oblique_runs" not in est and est["memory_bytes"] > 0
    est_o = api.estimate(lamellar([500.0, 600.0, 700.0], theta=60))
    assert est_o["runs"] == est_o["oblique_runs"] == 6 and est_o["memory_bytes"] > est["memory_bytes"]
    est_s = api.estimate(stack(CI(n_real=2.0), 100, pols=("s", "p"), angles=(0, 30, 60)))
    assert est_s["runs"] == 1 + 2 * 2 * 9  # normal incidence: one run serves s and p
    assert est_s["real_cells"][:2] == [1, 1]
    assert est_s["dt_fs"] > 0 and est_s["dft_stride"] >= 1 and est_s["steps_max"] % est_s["dft_stride"] == 0


def test_cell_plan_follows_the_reference_rules():
    opts = P.FDTDOptions()
    prob = stack(CI(n_real=2.0), 300)
    c = spectra.make_cell(prob, opts)
    d0 = 400.0 / (2.0 * 20)  # shortest wavelength over the densest index, 20 cells per wavelength
    assert c.spacing[2] == pytest.approx(d0) and c.n_cells == (1, 1)
    assert c.k_r < c.k_src < c.k_top < c.k_t < c.shape[2] - c.pml
    assert c.k_src == c.k_r + 2 and c.k_top - c.k_src >= 4
    assert c.k_t == c.k_top + math.ceil(300 / d0 - 1e-9) + 3
    assert c.origin[2] == pytest.approx(-c.k_top * d0)  # the stack's top (z = 0) on the node plane k_top
    thin = spectra.make_cell(stack(CI(n_real=2.0), 8.0), opts)
    assert thin.spacing[2] == pytest.approx(4.0)  # at least two cells across the thinnest layer
    obl = spectra.make_cell(stack(CI(n_real=2.0), 300, angles=(30.0,)), opts)
    assert obl.periods[0] == pytest.approx(d0 / 4)  # the invariant axis' nominal cell, narrower off-normal
    pc = spectra.make_cell(pillars(), opts)
    assert pc.n_cells[0] == pc.n_cells[1] == math.ceil(400 / (500 / (2.0 * 20)))
    assert pc.origin[0] == pytest.approx(-200.0 - pc.spacing[0])
    # the monitor gap: three decay lengths of the slowest evanescent order, within [0.25, 1.5] wavelengths
    gap_nm = (c.k_r - c.pml) * c.spacing[2]
    assert 0.25 * 800 - d0 <= gap_nm <= 1.5 * 800 + d0


def _brute_slot(prob, table, px, py, x, y, z):
    """The slot of one point by the definition: its layer by z, then the last shape that contains it or one of
    its periodic images."""
    tops = np.concatenate([[0.0], np.cumsum([lay.thickness_nm for lay in prob.layers])])
    if z < 0:
        return table.sup
    if z >= tops[-1]:
        return table.sub
    i = int(np.searchsorted(tops, z, side="right") - 1)
    bg, shapes = table.layers[i]
    one_d = prob.lattice_a2_nm is None
    out = bg
    for prof, s in shapes:
        for ix in range(-3, 4):
            for iy in ([0] if one_d else range(-3, 4)):
                xs, ys = x - ix * px, (0.0 if one_d else y - iy * py)
                if bool(spectra.inside_profile(prof, torch.tensor([xs]), torch.tensor([ys]))[0]):
                    out = s
    return out


def test_layer_slots_match_brute_force_with_periodic_images():
    prob = P.PeriodicProblem(
        lattice_a1_nm=(300, 0), lattice_a2_nm=(0, 250), superstrate=CI(n_real=1.0), substrate=CI(n_real=1.45),
        layers=[P.PatternedLayer(thickness_nm=100, background=CI(n_real=1.3), shapes=[
                    P.Structure2D(profile=P.Circle(center_nm=(140.0, 0.0), radius_nm=60), material=CI(n_real=2.0)),
                    P.Structure2D(profile=P.Rectangle(center_nm=(-20.0, 110.0), size_nm=(80.0, 60.0)),
                                  material=CI(n_real=2.5))]),
                P.PatternedLayer(thickness_nm=60, background=CI(n_real=1.6), shapes=[
                    P.Structure2D(profile=P.Polygon(vertices_nm=[(-150, -125), (0, 100), (120, -60)]),
                                  material=CI(n_real=2.0)),
                    P.Structure2D(profile=P.Ellipse(center_nm=(0, 0), semi_axes_nm=(90, 30), angle_deg=30),
                                  material=CI(n_real=3.0))])],
        wavelengths=[600.0])
    opts = P.FDTDOptions()
    cell = spectra.make_cell(prob, opts)
    table = spectra.MediaTable(prob, prob.wavelength_array(), [])
    assert table.sup == 0 and len(table.media) == 7  # superstrate, 1.3, 2.0, 2.5, 1.6, 3.0 (2.0 shared), substrate
    slots = spectra.LayerSlots(prob, table, cell)
    rng = np.random.default_rng(5)
    pts = np.column_stack([rng.uniform(-700, 700, 3000), rng.uniform(-600, 600, 3000), rng.uniform(-30, 200, 3000)])
    got = slots(torch.tensor(pts)).numpy()
    want = np.array([_brute_slot(prob, table, 300.0, 250.0, *p) for p in pts])
    assert np.array_equal(got, want)
    assert len(set(want.tolist())) == 7


def test_inside_profile_polygon_area_and_rectangle_equivalence():
    g = torch.linspace(-1, 1, 801, dtype=torch.float64)
    x, y = torch.meshgrid(g * 200, g * 200, indexing="ij")
    tri = P.Polygon(vertices_nm=[(-150, -100), (150, -100), (0, 160)])
    area = float(spectra.inside_profile(tri, x, y).double().mean()) * 400 ** 2
    assert abs(area - 0.5 * 300 * 260) < 0.01 * 0.5 * 300 * 260
    sq = P.Polygon(vertices_nm=[(-50.2, -30.2), (50.2, -30.2), (50.2, 30.2), (-50.2, 30.2)])
    rect = P.Rectangle(size_nm=(100.4, 60.4))
    assert torch.equal(spectra.inside_profile(sq, x, y), spectra.inside_profile(rect, x, y))
    ell = P.Ellipse(semi_axes_nm=(120, 40), angle_deg=90)
    tall = P.Ellipse(semi_axes_nm=(40, 120))
    assert torch.equal(spectra.inside_profile(ell, x, y), spectra.inside_profile(tall, x, y)) or \
        int((spectra.inside_profile(ell, x, y) ^ spectra.inside_profile(tall, x, y)).sum()) < 20


def test_rasterised_stack_has_the_layered_averaging_rules():
    """A 60 nm film of eps 4 between air and glass at 5 nm cells: E_x, E_y nodes inside the film take 1/4; the
    node planes on the two interfaces take 1 / <eps> over their cell (the tangential rule), the E_z planes, half a
    cell off the interfaces, are pure; above the stack 1, below 1/2.25."""
    prob = stack(CI(n_real=2.0), 60.0, spacing_nm=5.0)
    opts = P.FDTDOptions(spacing_nm=5.0)
    s = spectra.setup(prob, opts, "cpu")
    c = s["cell"]
    inv = s["pcell"].inv_eps[:, 1, 1, :]
    k0, k1 = c.k_top, c.k_top + 12
    for comp in (0, 1):
        col = inv[comp]
        assert np.allclose(col[2:k0], 1.0) and np.allclose(col[k0 + 1:k1], 0.25)
        assert np.allclose(col[k1 + 1:-1], 1 / 2.25)
        assert col[k0] == pytest.approx(1 / 2.5) and col[k1] == pytest.approx(1 / 3.125)
    colz = inv[2]
    assert np.allclose(colz[k0:k1], 0.25) and np.allclose(colz[2:k0], 1.0) and np.allclose(colz[k1:-2], 1 / 2.25)
    assert s["pcell"].offdiag is None and not s["pcell"].groups


def test_order_filter_keeps_the_propagating_orders():
    prob = lamellar([500.0, 900.0])
    cell = spectra.make_cell(prob, P.FDTDOptions())
    wl = prob.wavelength_array()
    orders, kpar2 = spectra.propagating_orders(prob, cell, wl, [(np.zeros(2), np.zeros(2))], 1.0, np.full(2, 2.25))
    ms = sorted(mn[0] for mn, _, _ in orders)
    assert ms == [-2, -1, 0, 1, 2]  # |m| 500 / 700 < 1.5 at 500 nm
    assert all(mn[1] == 0 for mn, _, _ in orders)
    nx = cell.n_cells[0]
    for (m, _), ii, _ in orders:
        assert np.fft.fftfreq(nx, 1 / nx)[ii] == m
    obl_k = 2 * np.pi / wl * math.sin(math.radians(30))
    orders_o, _ = spectra.propagating_orders(prob, cell, wl, [(obl_k, np.zeros(2))], 1.0, np.full(2, 2.25))
    assert sorted(mn[0] for mn, _, _ in orders_o) == [-2, -1, 0, 1]  # |0.5 + m 500 / 700| < 1.5: shifted by k_par
    st = stack(CI(n_real=2.0), 100)
    c2 = spectra.make_cell(st, P.FDTDOptions())
    o2, _ = spectra.propagating_orders(st, c2, st.wavelength_array(), [(np.zeros(9), np.zeros(9))], 1.0,
                                       np.full(9, 2.25))
    assert [mn for mn, _, _ in o2] == [(0, 0)]
