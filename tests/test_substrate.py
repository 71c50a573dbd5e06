"""Particles on a substrate at the solve level, on the CPU (no solve): the planar background the rasteriser starts
from is the meshing the source tables assume (the same arrays as tests/test_table_source.py's restatement, whose
leak tests pin it: swapping the averaging rules leaks); particles rasterised over it leave every node away from them
at its background value, and the vertex tensor of a metal particle on glass stays on the particle (the glass plane
is not grown into) and equals the oracle's restatement; the backend accepts what the HIP path runs over a
substrate and declines the rest with the reference's reasons (scattering.py:41-61)."""

import json

import numpy as np
import pytest
import torch

from lightsmith_amd import api, backend
from lightsmith_amd import problem as P
from lightsmith_amd.fdtd import scattering, voxel
from lightsmith_amd.fdtd.media import Medium, drude_lorentz
from oracle import tensor as otensor
from tests.test_table_source import DX, EPS, SHAPE, _layer_inv_eps


def sim_on(substrate, direction=(0.0, 0.0, -1.0), polarization=(1.0, 0.0, 0.0), material=None, z_c=22.0, **fdtd):
    """A 20 nm sphere (n = 2 unless given) centred z_c above glass (or ``substrate``) at z = 0, in air."""
    return P.Simulation(particles=[P.Particle(shape=P.Sphere(radius_nm=20.0), position_nm=(0.0, 0.0, z_c),
                                              material=material or P.ConstantIndex(n_real=2.0))],
                        medium_index=1.0, substrate=P.Substrate(material=substrate, interface_z_nm=0.0),
                        illumination=P.PlaneWave(direction=direction, polarization=polarization),
                        wavelengths=[450.0, 550.0, 650.0], options={"fdtd": {"spacing_nm": 2.5, **fdtd}})


GLASS = P.ConstantIndex(n_real=1.5)


def test_planar_background_is_the_meshing_the_tables_assume():
    for z_int in (37.0, 40.0, 42.0, 43.9):
        got = voxel.planar_background(SHAPE, np.zeros(3), DX, z_int, Medium(EPS), Medium(2.25))
        assert np.allclose(got, _layer_inv_eps(z_int, 2.25), rtol=0, atol=1e-15)
    inv = voxel.planar_background((3, 3, 6), np.zeros(3), 1.0, 2.3, Medium(1.0), Medium(2.25))
    assert inv[0][1, 1, 2] == pytest.approx(1 / (0.8 * 2.25 + 0.2))  # E_x cell [1.5, 2.5): 0.8 glass
    assert inv[2][1, 1, 2] == pytest.approx(0.3 / 2.25 + 0.7)  # E_z cell [2, 3): 0.3 glass, <1/eps>
    assert inv[0][1, 1, 1] == pytest.approx(1 / 2.25) and inv[0][1, 1, 3] == 1.0
    with pytest.raises(ValueError):
        voxel.planar_background((3, 3, 6), np.zeros(3), 1.0, 2.3, Medium(1.0), drude_lorentz(1.0, 9.0, 0.1))


def test_particles_over_the_background_leave_it_alone_and_the_tensor_stays_on_the_particle():
    shape, dx = (40, 40, 44), 1.0
    origin = -np.array([20.0, 20.0, 14.0])
    z_int, r, c = 0.0, 9.3, np.array([0.4, -0.3, 9.6])  # a Drude-Lorentz sphere resting on the glass
    metal = drude_lorentz(1.5, 9.0, 0.07, [(1.2, 2.8, 0.4)])
    media = [Medium(1.0), metal, Medium(2.25)]

    def t_slot(p):
        out = torch.where(((p - torch.as_tensor(c)) ** 2).sum(-1) <= r * r, 1, 0)
        return torch.where((out == 0) & (p[..., 2] < z_int), 2, out)

    def np_slot(p):
        out = np.where(np.sum((p - c) ** 2, axis=-1) <= r * r, 1, 0)
        return np.where((out == 0) & (p[..., 2] < z_int), 2, out)

    base = voxel.planar_background(shape, origin, dx, z_int, media[0], media[2])
    roi = (np.array([6, 6, 6]), np.array([34, 34, 38]))
    inv, groups, stats, od = voxel.voxelise(t_slot, media, shape, origin, dx, *roi, base=base, background=(0, 2))
    # every node whose cell is 2 cells clear of the sphere keeps its background value
    for comp in range(3):
        off = np.zeros(3)
        off[comp] = 0.5
        ii, jj, kk = np.meshgrid(*[np.arange(n) for n in shape], indexing="ij")
        pos = origin + (np.stack([ii, jj, kk], -1) + off) * dx
        far = np.linalg.norm(pos - c, axis=-1) > r + 2.0
        assert np.array_equal(inv[comp][far], base[comp][far])
    tm = [g for g in groups if hasattr(g, "slot_vert")]
    assert len(tm) == 1
    v = origin + tm[0].vert_idx3 * dx
    assert np.linalg.norm(v - c, axis=1).max() < r + 2.0  # not grown along the glass plane
    ref = otensor.vertex_tensor(np_slot, media, shape, origin, dx, *roi, samples=8, background=(0, 2))
    assert np.array_equal(tm[0].vert_idx3, ref.vert_idx3) and np.array_equal(tm[0].slot_vert, ref.slot_vert)
    assert np.array_equal(tm[0].node_idx3, ref.node_idx3) and np.array_equal(tm[0].frac, ref.frac)
    # without the background list the flood fill runs along the glass plane
    grown = voxel.vertex_tensor(t_slot, media, shape, origin, dx, *roi, samples=8)
    vg = origin + grown.vert_idx3 * dx
    assert np.linalg.norm(vg - c, axis=1).max() > r + 8.0 and len(grown.normal) > len(tm[0].normal) + 500


def test_substrate_support_and_refusals():
    B = backend.NativeFDTD
    assert B.supports(sim_on(GLASS))[0]
    assert B.supports(sim_on(GLASS, direction=(0.64, 0.0, -0.77), polarization=(0.77, 0.0, 0.64)))[0]
    assert B.supports(sim_on(GLASS, direction=(0.0, 0.5, 0.866), polarization=(1.0, 0.0, 0.0)))[0]  # through it
    ok, why = B.supports(sim_on(GLASS, direction=(0.995, 0.0, -0.0998), polarization=(0.0, 1.0, 0.0)))
    assert not ok and "graze" in why
    ok, why = B.supports(sim_on(P.ConstantIndex(n_real=0.2, n_imag=3.0)))
    assert not ok and "absorbing" in why
    ok, why = B.supports(sim_on(P.DrudeLorentz(eps_inf=2.2, poles=[P.LorentzPole(delta_eps=0.1, omega0_ev=9.0,
                                                                                  gamma_ev=0.0)])))
    assert not ok and "non-dispersive" in why
    back = P.parse_problem(json.loads(sim_on(GLASS).model_dump_json()))
    assert back.substrate.material == GLASS and back.substrate.interface_z_nm == 0.0


def test_substrate_plan_reaches_the_interface():
    est = api.estimate(sim_on(GLASS, z_c=60.0))
    lay = scattering.layout(sim_on(GLASS, z_c=60.0), P.FDTDOptions(spacing_nm=2.5), 2.5)
    z_lo = lay.origin[2] + lay.tfsf[0][2] * 2.5
    assert z_lo < -8.0  # the TFSF box reaches below the interface, with padding
    assert est["grid"] == list(lay.shape) and est["fits"]
    s = scattering.setup(sim_on(GLASS, z_c=60.0), P.FDTDOptions(spacing_nm=2.5), "cpu")
    assert s.substrate["incidence"] == "from above" and s.source.n_inc == pytest.approx(1.0)
    s2 = scattering.setup(sim_on(GLASS, direction=(0.0, 0.0, 1.0), z_c=30.0), P.FDTDOptions(spacing_nm=2.5), "cpu")
    assert s2.substrate["incidence"] == "through the substrate" and s2.source.n_inc == pytest.approx(1.5)
    # the rasterised glass far from the sphere is the planar background
    k = np.arange(s.layout.shape[2])
    z = s.layout.origin[2] + k * 2.5
    col = s.inv_eps[0][2, 2, :]
    assert np.allclose(col[z < -2.0], 1 / 2.25) and np.allclose(col[z > 2.0], 1.0)


def test_near_field_slab_covers_the_plane():
    sim = P.Simulation(particles=[P.Particle(shape=P.Sphere(radius_nm=10), material=P.ConstantIndex(n_real=2.0))],
                       wavelengths=[500.0], options={"fdtd": {"spacing_nm": 4.0}},
                       outputs=P.Outputs(near_field=P.NearFieldRequest(plane="yz", offset_nm=2.0, half_width_nm=30,
                                                                       points=5, wavelengths_nm=[500.0, 600.0])))
    lay = scattering.layout(sim, P.FDTDOptions(spacing_nm=4.0), 4.0)
    near = scattering.near_slab(sim, lay)
    x = lay.origin[0] + near["ranges"][0] * 4.0
    y = lay.origin[1] + near["ranges"][1] * 4.0
    assert x.min() <= 2.0 - 4.0 and x.max() >= 2.0 + 4.0  # a cell either side of the plane x = 2
    assert y.min() <= -34.0 and y.max() >= 34.0
    assert len(near["idx"]) == 3 * np.prod([len(r) for r in near["ranges"]])
    assert np.allclose(near["omega"], 2 * np.pi / np.array([500.0, 600.0]))
    assert scattering.band_wavelengths(sim).tolist() == [500.0, 500.0, 600.0]
