#!/usr/bin/env python
"""Throughput benchmark of the FDTD time-stepping hot path on one MI355X.

Workload: the plain-FDTD throughput configuration (lightsmith_amd/fdtd/workloads.py): 320^3 cells, 16-cell
CPML, 200 seeded dielectric spheres, 10 % of the non-CPML cells D-form 3-pole Lorentz ADE, a point dipole,
a 16-frequency DFT plane; float32 fields. A step is one FDTD time step over the whole grid (H update, E
update with CPML and ADE, dipole, and one DFT sample every dft_stride steps), replayed as HIP graphs of one
DFT chunk each, inputs resident in HBM.

Output: one JSON line on stdout. value = cells x steps / wall time of the timed steps (Gcell-updates/s).
roofline: the bulk update (H kernel + E kernel) against the HBM roofline: algorithmic bytes per step
(fields, coefficients, CPML psi, ADE state, each read and written once) / the two kernels' mean durations,
measured with HIP events around each launch on the launch stream over a second pass of the same steps.
cpu_baseline: the oracle restated in C (float64, OpenMP over the host's cores) on a bounded sample of the same
scene; the FP64 NumPy oracle (one thread) alongside.
batched: the oblique sweep (64 uneven Bloch-periodic runs) stepped as one autotuned Batch against the same runs
stepped one after another on the plain engine; aggregate Gcell-updates/s of both and their ratio.
adjoint: time of dL/drho (forward + adjoint run) over one forward run on the 256x192x96 adjoint scene.
au_sphere: the plain Au sphere (97^3, vertex tensor, table-fed source, 81-wavelength flux boxes), 4000 steps.
batched_adjoint: 16 dispersive inverse-design instances, one batched gradient against 16 sequential ones.
"""

from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

METRIC = "Gcell-updates/s + % HBM roofline; batched and adjoint vs plain GPU FDTD"
HBM_PEAK_GBS = 8000.0
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic_throughput_320.json")


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def algorithmic_bytes(eng, scene) -> dict:
    """Bytes per step that the hot path must move at minimum (field, coefficient, psi, ADE terms of
    the per-step formula; plus the dipole's and the DFT plane's share)."""
    b = eng.bytes_per_step()
    s = 8 if eng.complex else 4
    n_src = 2 * 1  # dipole nodes: target read + write, one table tap
    src = n_src * (4 + 2 * s)
    e_idx = scene["monitor"][0]
    mon = len(scene["omega"]) * len(e_idx) * 2 * 2 * 8 / eng.stride  # acc_e + acc_h complex64 r/w per sample
    b["sources"] = src
    b["monitors"] = mon
    b["step_total"] = b["total"] + src + mon
    return b


def cpu_baseline(scene, steps: int, c_steps: int) -> dict:
    """The oracle on the same scene: the C restatement (oracle/cyee.c, float64, OpenMP over the host's cores)
    timed over `c_steps` steps, and the FP64 NumPy oracle (one thread) over `steps` steps, each after two warm-up
    steps. The C figure is the baseline reported; the NumPy one rides along."""
    from oracle import cyee
    from oracle import fdtd as ofd
    from oracle import media as omed
    from oracle import sources as osrc

    m = scene["medium"]
    med = omed.Medium(m.eps_inf, tuple(omed.Pole(p.sigma, p.omega0, p.gamma) for p in m.poles))
    idx = scene["disp_idx"]
    groups = [ofd.DispersiveGroup(idx=idx, media=[med], frac=np.ones((1, len(idx))), w=None)]
    src = osrc.PointDipole(shape=scene["shape"], origin=np.zeros(3), dx=scene["dx"], dt=scene["dt"],
                           pos=scene["dipole_pos"], pdir=scene["dipole_dir"], inv_eps=scene["inv_eps"],
                           omega_c=scene["omega_c"], omega_half=scene["omega_half"],
                           max_steps=max(steps, c_steps) + 10)
    mon = ofd.DFTMonitor(scene["omega"], *scene["monitor"], energy=1)
    p = scene["pml"]
    pml = ofd.PMLSpec(cells=((p, p),) * 3, alpha_max=0.05 * float(np.min(scene["omega"])), eps_ref=1.0)
    eng = ofd.Engine(shape=scene["shape"], dx=scene["dx"], dt=scene["dt"], inv_eps=scene["inv_eps"], groups=groups,
                     pml=pml, sources=[src], monitors=[mon], dft_stride=scene["stride"])
    cells = scene["n_cells"]
    threads = int(os.environ.get("OMP_NUM_THREADS") or os.cpu_count())
    c = cyee.CYee(eng)
    c.run(2)
    t0 = time.perf_counter()
    c.run(c_steps)
    c_dt = time.perf_counter() - t0
    for _ in range(2):
        eng.step()
    t0 = time.perf_counter()
    for _ in range(steps):
        eng.step()
    np_dt = time.perf_counter() - t0
    return {"value": cells * c_steps / c_dt / 1e9, "unit": "Gcell-updates/s", "cores": threads, "kind": "port",
            "sample": f"{c_steps} steps of the same {scene['shape'][0]}^3 scene (H, E, CPML, ADE, dipole, DFT samples "
                      f"every {scene['stride']} steps) after 2 warm-up steps; the oracle restated in C "
                      f"(oracle/cyee.c: float64, OpenMP over {threads} threads, fields equal to the NumPy oracle's)",
            "seconds": c_dt,
            "numpy_one_thread": {"value": cells * steps / np_dt / 1e9, "cores": 1, "steps": steps,
                                 "seconds": np_dt}}


def batched_leg(dev, steps: int, n_inst: int = 64, tune: bool = True) -> dict:
    """The oblique sweep (workloads.oblique_sweep_scene: n_inst Bloch-periodic metasurface runs of uneven
    cells, complex64 fields for oblique incidence, real at normal incidence) stepped two ways over the same
    `steps` steps per run: sequentially, each run alone with its own HIP graph per DFT chunk (the plain
    engine), and as one Batch (one launch per half-step over every run's tiles, one HIP graph per chunk,
    autotuned schedule). Returns both aggregate rates (cells x steps / time over the timed region)."""
    import torch

    from lightsmith_amd.fdtd import workloads
    from lightsmith_amd.fdtd.batch import Batch
    from lightsmith_amd.fdtd.engine import filling_tiling

    t0 = time.perf_counter()
    runs = workloads.oblique_sweep_scene(n_inst=n_inst)
    stride = workloads.common_stride(runs)
    K = workloads.steps_for(steps, stride)
    engs = [workloads.oblique_engine(r, dev, max_steps=3 * K + 8 * stride, stride=stride) for r in runs]
    for e in engs:  # each sequential run at the tiling the periodic driver gives a grid of its size
        e.set_tiling(*filling_tiling([e.shape]))
    torch.cuda.synchronize()
    cells = sum(e.N for e in engs)
    log(f"oblique sweep: {len(engs)} runs, {cells / 1e6:.1f} M cells, stride {stride}, set-up "
        f"{time.perf_counter() - t0:.1f} s")

    # sequential plain runs
    seq = 0.0
    for e in engs:
        run, _ = e.make_runner(True)
        run()  # eager warm-up chunk + capture
        run()
        torch.cuda.synchronize()
        a = time.perf_counter()
        for _ in range(K // stride):
            run()
        torch.cuda.synchronize()
        seq += time.perf_counter() - a
        e.release()
    seq_rate = cells * K / seq / 1e9
    log(f"sequential: {seq * 1e3:.1f} ms for {K} steps of every run ({seq_rate:.2f} Gcell-updates/s)")

    b = Batch(engs)
    top = []
    if tune:
        a = time.perf_counter()
        res = b.autotune(steps=6, warmup=2)
        top = [[round(ms, 4), list(c)] for ms, c in res[:5]]
        log(f"autotune: {len(res)} schedules in {time.perf_counter() - a:.1f} s; best {res[0][1]} "
            f"({res[0][0]:.3f} ms per step of bulk kernels), worst {res[-1][0]:.3f} ms")
    run, graphs = b.make_runner(True)
    run()
    run()
    torch.cuda.synchronize()
    a = time.perf_counter()
    for _ in range(K // stride):
        run()
    torch.cuda.synchronize()
    bt = time.perf_counter() - a
    b.release()
    bat_rate = cells * K / bt / 1e9
    h_ms, e_ms = b.step_timed(2 * stride)
    kern_s = (h_ms + e_ms) / (2 * stride) / 1e3
    bytes_step = sum(e.bytes_per_step()["total"] for e in engs)
    achieved = bytes_step / kern_s / 1e9
    log(f"batched: {bt * 1e3:.1f} ms ({bat_rate:.2f} Gcell-updates/s, {bat_rate / seq_rate:.2f}x sequential); "
        f"bulk kernels {kern_s * 1e3:.3f} ms per step, {achieved:.0f} GB/s algorithmic")
    finite = all(np.isfinite(float(e.energy[0])) for e in engs)
    b.close()
    for e in engs:
        e.close()
    return {"workload": f"oblique sweep: {len(engs)} Bloch-periodic metasurface runs (8 angles 0-85 deg x 8 "
                        "wavelengths 500-1500 nm, s/p alternating), cells 48-96 x 48-96 x 256 + ghosts, pillar "
                        "layer eps 1/12.1 on eps 2.1, CPML in z, Bloch plane source (TFSF line at normal "
                        "incidence), R/T/incident DFT monitors",
            "runs": len(engs), "cells": int(cells), "steps_per_run": K, "dft_stride": stride,
            "sequential_gcell_per_s": round(seq_rate, 3), "batched_gcell_per_s": round(bat_rate, 3),
            "speedup": round(bat_rate / seq_rate, 3), "schedule": list(b.schedule), "hip_graphs": graphs,
            "roofline": {"achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": round(achieved / HBM_PEAK_GBS, 4), "bulk_kernel_ms_per_step": round(kern_s * 1e3, 4),
                         "algorithmic_bytes_per_step": int(bytes_step)},
            "autotune_top5": top, "finite": finite}


def adjoint_leg(dev, repeats: int = 2) -> dict:
    """Cost of dL/drho by the adjoint (forward run + adjoint run, lightsmith_amd.fdtd.adjoint) against one forward
    run, on the adjoint scene (workloads.adjoint_scene: 256x192x96 cells, 80x80x11 design region, 3000 steps):
    with the stored tape and with checkpoint recompute every 100 steps. Wall time between device
    synchronisations, best of `repeats`."""
    import torch

    from lightsmith_amd.fdtd import workloads
    from lightsmith_amd.fdtd.adjoint import monitor_spectra

    t0 = time.perf_counter()
    scene = workloads.adjoint_scene()
    prob = workloads.adjoint_problem(scene, dev)
    rho0 = torch.tensor(scene["rho"], dtype=torch.float64, device=dev)
    prob.forward(rho0)  # warm-up (captures the chunk graph)
    torch.cuda.synchronize()
    log(f"adjoint scene: {scene['shape']} cells, design {prob.design.shape}, {prob.n_steps} steps, tape "
        f"{prob.tape.numel() * 4 / 1e9:.2f} GB, set-up {time.perf_counter() - t0:.1f} s")

    def timed(fn):
        best = float("inf")
        for _ in range(repeats):
            torch.cuda.synchronize()
            a = time.perf_counter()
            out = fn()
            torch.cuda.synchronize()
            best = min(best, time.perf_counter() - a)
        return best, out

    t_fwd, _ = timed(lambda: prob.forward(rho0))

    def grad():
        rho = rho0.clone().requires_grad_(True)
        acc_e, _ = monitor_spectra(rho, prob)
        loss = workloads.mode_power(acc_e, scene["w"])
        loss.backward()
        return rho.grad, float(loss.detach())

    t_grad, (g, loss) = timed(grad)
    prob.use_checkpoints(True)
    t_ck, (g_ck, _) = timed(grad)
    prob.use_checkpoints(False)
    same = bool(torch.equal(g, g_ck))
    log(f"adjoint: forward {t_fwd * 1e3:.1f} ms, forward + adjoint {t_grad * 1e3:.1f} ms "
        f"({t_grad / t_fwd:.2f}x), with checkpoint recompute {t_ck * 1e3:.1f} ms ({t_ck / t_fwd:.2f}x)")
    cells = scene["n_cells"]
    prob.close()
    return {"workload": "adjoint inverse design: 256x192x96 cells at 20 nm, Si strip waveguide (eps 3.48^2) in "
                        "oxide (1.44^2), 80x80x11 design region rho~U[0,1] (seed 4), mode-injection source and "
                        "mode-power objective at 1550 nm, 16-cell CPML, 3000 steps, float32",
            "forward_ms": round(t_fwd * 1e3, 2), "gradient_ms": round(t_grad * 1e3, 2),
            "gradient_vs_forward": round(t_grad / t_fwd, 3),
            "gradient_checkpointed_ms": round(t_ck * 1e3, 2),
            "gradient_checkpointed_vs_forward": round(t_ck / t_fwd, 3), "checkpoint_every": prob.K,
            "checkpointed_equals_stored": same, "forward_gcell_per_s": round(cells * prob.n_steps / t_fwd / 1e9, 3),
            "loss": loss, "grad_norm": float(torch.linalg.vector_norm(g))}


def read_traffic():
    if not os.path.exists(TRAFFIC_FILE):
        return None
    with open(TRAFFIC_FILE) as f:
        return json.load(f)


def au_sphere_leg(dev) -> dict:
    """The plain Au sphere (workloads.au_sphere_sim: 97^3 nodes at 2 nm, 12-cell CPML, Drude + 2 Lorentz gold with
    the conformal vertex tensor, table-fed plane wave 400-800 nm, 81-wavelength flux boxes, DFT every 100 steps):
    4000 steps through HIP graphs, the whole step timed (bulk kernels, table source, vertex tensor, pure ADE, DFT
    of 71,664 flux points x 81 wavelengths); the bulk kernels timed alone with HIP events for the roofline."""
    import torch

    from lightsmith_amd.fdtd import workloads

    t0 = time.perf_counter()
    sim, opts, scene = workloads.au_sphere_scene(dev, dft_stride=100)
    eng = scene.engine(dev)
    t_set = time.perf_counter() - t0
    # a 97-cell row fills one 64-cell wavefront row and a third of a second, or three 32-cell rows and one lane of
    # a fourth, and 0.91 M cells fill the GPU only with short z marches: pick the tiling by timing a few (each
    # engine steps alone, so any compiled shape is exact)
    tilings = [(1, 4, 64, 8), (1, 4, 32, 8), (1, 4, 32, 4), (1, 4, 64, 4), (1, 4, 64, 3), (1, 4, 64, 2),
               (1, 4, 64, 1), (1, 8, 64, 2)]
    times = {}
    for t in tilings:
        eng.set_tiling(*t)
        eng.steps(20)
        torch.cuda.synchronize()
        a = time.perf_counter()
        eng.steps(200)
        torch.cuda.synchronize()
        times[t] = (time.perf_counter() - a) / 200
    best = min(times, key=times.get)
    for tt in eng.state_tensors():
        tt.zero_()
    eng.n_host = 0
    eng.set_tiling(*best)
    steps = 4000
    run, used = eng.make_runner(True)
    run()
    torch.cuda.synchronize()
    a = time.perf_counter()
    for _ in range(steps // eng.stride - 1):
        run()
    torch.cuda.synchronize()
    wall = time.perf_counter() - a
    timed = steps - eng.stride
    eng.release()
    ms = (ctypes.c_double * 2)(0.0, 0.0)
    stream = ctypes.c_void_p(torch.cuda.current_stream(eng._dev).cuda_stream)
    from lightsmith_amd import _lib
    _lib.check(eng.lib.lsf_step_timed(eng._h, 400, stream, ms), "lsf_step_timed")
    kern_s = (ms[0] + ms[1]) / 400 / 1e3
    b = eng.bytes_per_step()
    n = eng.N
    achieved = b["total"] / kern_s / 1e9
    out = {"workload": "plain Au sphere: r = 40 nm Drude + 2 Lorentz gold in water, 97^3 nodes at 2 nm (12-cell CPML), "
                       "conformal vertex tensor, table-fed plane wave 400-800 nm, 81-wavelength flux boxes, "
                       f"{steps} steps, DFT every 100, float32, HIP graphs",
           "cells": n, "steps": timed, "ms_per_step": round(wall / timed * 1e3, 4), "tiling": list(best),
           "tiling_trial_us_per_step": {",".join(map(str, t)): round(v * 1e6, 1) for t, v in times.items()},
           "gcell_per_s": round(n * timed / wall / 1e9, 3), "hip_graphs": used, "setup_s": round(t_set, 2),
           "voxelise_s": round(scene.voxelise_s, 3), "tensor_vertices": scene.stats["tensor_vertices"],
           "dispersive_cells": scene.stats["dispersive_cells"],
           "bulk_kernels_us": {"h": round(ms[0] / 400 * 1e3, 2), "e": round(ms[1] / 400 * 1e3, 2)},
           "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": round(achieved / HBM_PEAK_GBS, 4), "algorithmic_bytes_per_step": int(b["total"])}}
    log(f"Au sphere: {out['gcell_per_s']} Gcell-updates/s ({out['ms_per_step'] * 1e3:.1f} us per step), bulk kernels "
        f"{achieved:.0f} GB/s, set-up {t_set:.1f} s")
    eng.close()
    return out


def batched_adjoint_leg(dev, steps: int = 2500) -> dict:
    """Batched forward + adjoint with dispersion (workloads.batched_adjoint_scenes: 16 uneven waveguide instances,
    each with a density design region and a Drude + 3 Lorentz silver cube, checkpoint recompute every 100 steps):
    one batched gradient of all instances against the same 16 gradients computed one instance after another."""
    import torch

    from lightsmith_amd.fdtd import workloads
    from lightsmith_amd.fdtd.adjoint import BatchedAdjoint, batched_monitor_spectra, monitor_spectra

    scenes = workloads.batched_adjoint_scenes(n_steps=steps)
    probs = [workloads.adjoint_problem(sc, dev, checkpoint_every=100, tape_budget=1) for sc in scenes]
    cells = sum(sc["n_cells"] for sc in scenes)

    def seq():
        gs = []
        for p, sc in zip(probs, scenes):
            rho = torch.tensor(sc["rho"], dtype=torch.float64, device=dev, requires_grad=True)
            acc_e, _ = monitor_spectra(rho, p)
            workloads.mode_power(acc_e, sc["w"]).backward()
            gs.append(rho.grad)
        return gs

    seq()  # warm-up: each instance's chunk graph
    torch.cuda.synchronize()
    a = time.perf_counter()
    g_seq = seq()
    torch.cuda.synchronize()
    t_seq = time.perf_counter() - a
    for p in probs:
        p.release_graphs()
    batch = BatchedAdjoint(probs)

    def bat():
        rhos = [torch.tensor(sc["rho"], dtype=torch.float64, device=dev, requires_grad=True) for sc in scenes]
        outs = batched_monitor_spectra(rhos, batch)
        sum(workloads.mode_power(e, sc["w"]) for (e, _), sc in zip(outs, scenes)).backward()
        return [r.grad for r in rhos]

    bat()
    torch.cuda.synchronize()
    a = time.perf_counter()
    g_bat = bat()
    torch.cuda.synchronize()
    t_bat = time.perf_counter() - a
    same = all(bool(torch.equal(x, y)) for x, y in zip(g_seq, g_bat))
    batch.close()
    for p in probs:
        p.close()
    log(f"batched adjoint: 16 gradients {t_bat * 1e3:.0f} ms batched, {t_seq * 1e3:.0f} ms one after another "
        f"({t_seq / t_bat:.2f}x); identical {same}")
    return {"workload": "batched adjoint with dispersion: 16 uneven waveguide instances (128-256 x 96-192 x 64-96 "
                        "cells, seed 6), density design region + Drude + 3 Lorentz silver cube each, "
                        f"{steps} steps, checkpoint recompute every 100 steps, float32",
            "instances": len(scenes), "cells": int(cells), "steps": steps,
            "batched_ms": round(t_bat * 1e3, 1), "sequential_ms": round(t_seq * 1e3, 1),
            "speedup": round(t_seq / t_bat, 3), "batched_equals_sequential": same,
            "gradient_cell_updates_per_s": {"batched": round(cells * steps / t_bat / 1e9, 3),
                                            "sequential": round(cells * steps / t_seq / 1e9, 3)}}


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--n", type=int, default=320, help="grid size per axis (the workload is 320)")
    ap.add_argument("--cpu-steps", type=int, default=3, help="timed steps of the NumPy CPU oracle (0: skip the CPU legs)")
    ap.add_argument("--cpu-c-steps", type=int, default=300,
                    help="timed steps of the C (OpenMP) CPU oracle (about 10 s on 16 threads)")
    ap.add_argument("--tiling", default=None, help="rows_per_lane,waves,x_cells,z_chunk")
    ap.add_argument("--single-pass", choices=["auto", "on", "off"], default="auto",
                    help="step with the single-pass kernel (on), the H and E kernels (off), or the engine default")
    ap.add_argument("--no-roofline-pass", action="store_true")
    ap.add_argument("--eager", action="store_true", help="launch each chunk's kernels directly instead of a HIP graph")
    ap.add_argument("--batched-steps", type=int, default=300,
                    help="steps per run of the batched-vs-sequential oblique sweep (0: skip)")
    ap.add_argument("--batched-only", action="store_true", help="run only the oblique sweep comparison")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint cost measurement")
    ap.add_argument("--no-au-sphere", action="store_true", help="skip the plain Au sphere leg")
    ap.add_argument("--au-sphere-only", action="store_true", help="run only the plain Au sphere leg")
    ap.add_argument("--batched-adjoint-only", action="store_true", help="run only the batched dispersive adjoint leg")
    ap.add_argument("--no-batched-adjoint", action="store_true", help="skip the batched dispersive adjoint leg")
    ap.add_argument("--adjoint-only", action="store_true", help="run only the adjoint cost measurement")
    ap.add_argument("--no-tune", action="store_true", help="batched sweep: keep the default schedule")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain device copy of 1 GiB (for scaling rocprof FETCH/WRITE counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU only (one process, one MI355X); --gpus above 1 is not "
              "supported", file=sys.stderr)
        return 2

    import torch

    from lightsmith_amd import _lib
    from lightsmith_amd.fdtd import workloads

    if not torch.cuda.is_available():
        print("bench.py: no HIP device visible", file=sys.stderr)
        return 1
    dev = torch.device("cuda:0")
    if args.copy_calibration:
        a = torch.ones(1 << 28, dtype=torch.float32, device=dev)
        b = torch.empty_like(a)
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        log("copy calibration: 5 copies of 1 GiB")
        return 0

    if args.batched_only:
        print(json.dumps(batched_leg(dev, args.batched_steps, tune=not args.no_tune)), flush=True)
        return 0
    if args.adjoint_only:
        print(json.dumps(adjoint_leg(dev)), flush=True)
        return 0
    if args.au_sphere_only:
        print(json.dumps(au_sphere_leg(dev)), flush=True)
        return 0
    if args.batched_adjoint_only:
        print(json.dumps(batched_adjoint_leg(dev)), flush=True)
        return 0

    t_setup = time.perf_counter()
    scene = workloads.throughput_scene(n=args.n)
    stride = scene["stride"]
    K = workloads.steps_for(args.steps, stride)
    W = workloads.steps_for(max(args.warmup, stride), stride)
    tiling = tuple(int(v) for v in args.tiling.split(",")) if args.tiling else None
    fused = {"auto": None, "on": True, "off": False}[args.single_pass]
    eng = workloads.throughput_engine(scene, dev, max_steps=W + 2 * K + 4 * stride, tiling=tiling, fused=fused)
    torch.cuda.synchronize()
    log(f"set-up {time.perf_counter() - t_setup:.1f} s: {scene['shape']} cells, dt {scene['dt']:.4f} nm/c, "
        f"stride {stride}, {eng.n_ade} D-form nodes, {eng.n_psi} psi values, material {eng.material_mode}")

    run, used = eng.make_runner(not args.eager)
    for _ in range(W // stride):
        run()
    torch.cuda.synchronize()
    log(f"warm-up {W} steps done")

    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    start.record()
    for _ in range(K // stride):
        run()
    stop.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_ev = start.elapsed_time(stop)
    eng.release()
    elapsed = max(wall, ms_ev / 1e3)
    graphs = used
    value = scene["n_cells"] * K / elapsed / 1e9
    log(f"timed {K} steps: {elapsed * 1e3:.1f} ms ({value:.2f} Gcell-updates/s)")
    energy = float(eng.energy[0])
    if not np.isfinite(energy):
        print("bench.py: fields diverged", file=sys.stderr)
        return 1

    bytes_ = algorithmic_bytes(eng, scene)
    roofline = None
    if not args.no_roofline_pass:
        ms = (ctypes.c_double * 2)(0.0, 0.0)
        stream = ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)
        for _ in range(K // stride):
            _lib.check(eng.lib.lsf_step_timed(eng._h, stride, stream, ms), "lsf_step_timed")
            _lib.check(eng.lib.lsf_sample(eng._h, stream), "lsf_sample")
        torch.cuda.synchronize()
        h_ms, e_ms = ms[0] / K, ms[1] / K
        kern_s = (h_ms + e_ms) / 1e3
        achieved = bytes_["total"] / kern_s / 1e9
        traffic = read_traffic()
        traffic_src = None
        if traffic is not None:
            # the counters cannot be read inside this process; the file records the kernel durations of the
            # profiled run, so a traffic figure from an older build of the kernels shows up as stale here
            then = traffic.get("kernel_us") or {}
            now = {"h": h_ms * 1e3, "e": e_ms * 1e3}
            stale = not then or any(abs(now[k] - then.get(k, 0.0)) > 0.05 * now[k] for k in now)
            traffic_src = {"file": os.path.relpath(TRAFFIC_FILE, ROOT), "kernel_us_profiled": then,
                           "stale": bool(stale), "command": traffic.get("command")}
        roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(achieved / HBM_PEAK_GBS, 4),
                    "traffic": None if traffic is None else traffic.get("bytes_per_step"),
                    "traffic_source": traffic_src,
                    "kernel": ("k_step_fused (one time step's bulk update in a single pass)" if h_ms == 0.0 else
                               "k_update_h + k_update_e (one time step's bulk update)"),
                    "algorithmic_bytes_per_step": int(bytes_["total"]),
                    "kernel_us": {"h": round(h_ms * 1e3, 2), "e": round(e_ms * 1e3, 2)},
                    "step_achieved_GBs": round(bytes_["step_total"] / elapsed * K / 1e9, 1)}
        log(f"kernels: H {h_ms * 1e3:.1f} us, E {e_ms * 1e3:.1f} us per step; {achieved:.0f} GB/s algorithmic "
            f"({achieved / HBM_PEAK_GBS:.1%} of {HBM_PEAK_GBS:.0f})")

    cpu = None
    if args.cpu_steps > 0:
        log(f"CPU baseline: {args.cpu_c_steps} steps of the C oracle, {args.cpu_steps} of the NumPy oracle ...")
        cpu = cpu_baseline(scene, args.cpu_steps, args.cpu_c_steps)
        log(f"CPU baseline {cpu['value'] * 1e3:.2f} Mcell-updates/s on {cpu['cores']} threads "
            f"({cpu['seconds']:.1f} s); NumPy one thread {cpu['numpy_one_thread']['value'] * 1e3:.2f}")

    eng.close()
    batched = None
    if args.batched_steps > 0:
        batched = batched_leg(dev, args.batched_steps, tune=not args.no_tune)
    adjoint = None if args.no_adjoint else adjoint_leg(dev)
    au = None if args.no_au_sphere else au_sphere_leg(dev)
    badj = None if args.no_batched_adjoint else batched_adjoint_leg(dev)

    out = {"metric": METRIC, "value": round(value, 3), "unit": "Gcell-updates/s", "n_gpus": 1, "steps": K,
           "warmup": W, "steps_requested": args.steps, "warmup_requested": args.warmup,
           "steps_note": (None if (K, W) == (args.steps, args.warmup) else
                          f"steps and warm-up rounded up to whole DFT chunks of {stride} steps (one HIP graph "
                          "replay = one chunk = stride steps + one DFT sample)"), "ms_per_step": round(elapsed * 1e3 / K, 4), "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f32", "data": "synthetic",
           "config": {"workload": f"plain 3D FDTD throughput: {args.n}^3 cells, 16-cell CPML, 200 seeded spheres "
                                  "(eps 1-12), 10% of non-CPML cells 3-pole Lorentz ADE, z point dipole at 1550 nm, "
                                  "16-wavelength DFT plane, float32 fields, HIP graphs per DFT chunk",
                      "cells": scene["n_cells"], "dx_nm": scene["dx"], "dt_nm_per_c": scene["dt"],
                      "dft_stride": stride, "dispersive_nodes": int(eng.n_ade), "psi_values": int(eng.n_psi),
                      "material": eng.material_mode, "tiling": list(eng.tiling if tiling is None else tiling),
                      "tiling_note": "rows per lane, wavefronts per block, x cells per row, z planes per tile; the E "
                                     "kernel of an engine with pole state marches twice the z planes",
                      "hip_graphs": graphs, "parallelism": "single GPU"},
           "roofline": roofline, "cpu_baseline": cpu, "batched": batched, "adjoint": adjoint, "au_sphere": au,
           "batched_adjoint": badj}
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
