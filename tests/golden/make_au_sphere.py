"""Writes tests/golden/au_sphere_4000.npz from the FP64 oracle (``python -m tests.golden.make_au_sphere``, about
six minutes on one core).

The plain Au sphere (lightsmith_amd.fdtd.workloads.au_sphere_sim: r = 40 nm, Drude + 2 Lorentz Au in water,
2 nm cells, 97^3 nodes with a 12-cell CPML, table-fed plane wave 400-800 nm, 81 wavelengths), run for 4000 steps
(DFT every 100 steps) by the oracle on the inputs the product's set-up builds (voxelised materials with the vertex
tensor, source tables, flux boxes). Stored: the incident reference spectra, the net fluxes of the scattered- and
total-field boxes per wavelength, the derived cross sections, the monitor energies, and the DFT spectra at 64 seeded
points of each box (E and face-averaged H) -- enough to check the HIP engine's 4000-step spectra without the
86 MB of full monitor arrays. The inputs are rebuilt by the same set-up code in the test.

The same run with the oracle in float32 (the reference's algorithm in float32, tables included) is stored as the
float32 floor of each compared quantity (``f32_*``): the relative L2 distance of the float32 run from the float64
one. The HIP engine's float32 results are held to that floor where it exceeds 1e-5.
"""

from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

from tests import cases  # noqa: E402

STEPS, STRIDE, N_PICK = 4000, 100, 64
OUT = os.path.join(HERE, "au_sphere_4000.npz")


def picks(n: int, seed: int) -> np.ndarray:
    return np.sort(np.random.default_rng(seed).choice(n, size=N_PICK, replace=False))


def run(scene, dtype):
    from lightsmith_amd.fdtd import scattering
    from lightsmith_amd.fdtd.run import flux
    eng = cases.oracle_scattering_engine(scene, dtype=dtype)
    for q in range(STEPS // STRIDE):
        eng.chunk()
        print(f"{np.dtype(dtype).name}: chunk {q + 1}/{STEPS // STRIDE}", flush=True)
    (es, hs), (et, ht), (e_ref, _) = (m.results() for m in eng.monitors)
    c_sca, c_abs, c_ext = scattering.cross_sections(scene, e_ref, es, hs, et, ht)
    return dict(e_ref=e_ref, flux_sca=flux(es, hs, scene.sca[3]), flux_tot=flux(et, ht, scene.tot[3]), c_sca=c_sca,
                c_abs=c_abs, c_ext=c_ext, energy=np.asarray(eng.energy, dtype=float), es=es, hs=hs, et=et, ht=ht)


def rel(a, b) -> float:
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


def main():
    from lightsmith_amd.fdtd import workloads
    sim, opts, scene = workloads.au_sphere_scene("cpu", dft_stride=STRIDE)
    r64 = run(scene, np.float64)
    r32 = run(scene, np.float32)
    ps, pt = picks(r64["es"].shape[1], 1), picks(r64["et"].shape[1], 2)
    sub = lambda r: dict(es=r["es"][:, ps], hs=r["hs"][:, ps], et=r["et"][:, pt], ht=r["ht"][:, pt])  # noqa: E731
    out64 = {k: r64[k] for k in ("e_ref", "flux_sca", "flux_tot", "c_sca", "c_abs", "c_ext", "energy")} | sub(r64)
    out32 = {k: r32[k] for k in ("e_ref", "flux_sca", "flux_tot", "c_ext")} | sub(r32)
    floor = {f"f32_{k}": rel(v, out64[k]) for k, v in out32.items()}
    np.savez_compressed(OUT, steps=STEPS, stride=STRIDE, pick_sca=ps, pick_tot=pt, **out64, **floor)
    print("wrote", OUT, {k: f"{v:.2e}" for k, v in floor.items()})


if __name__ == "__main__":
    main()
