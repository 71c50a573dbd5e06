"""Run control and flux geometry around the engine (setup.py:16-17, 432-468, 564-566, 722-773;
periodic.py:335-363; progress.py:25-42).

``run_until_decayed`` drives any engine with the reference's contract (``make_runner``, ``energy``,
``stride``, ``release``): one 2-float readback per chunk, divergence -> FloatingPointError, decay
shut-off over a sliding window, progress events through a context-local callback that may raise
``Cancelled``.
"""

from __future__ import annotations

import contextvars
import math
from collections import deque
from contextlib import contextmanager
from dataclasses import dataclass

import numpy as np

from .engine import filling_tiling, flat_index

FS = 299.792458  # nm/c per femtosecond (setup.py:16)
DIVERGED_GROWTH = 1e8  # monitor energy / its peak while the pulse passed beyond which a run has diverged (setup.py:17)

_callback: contextvars.ContextVar = contextvars.ContextVar("lightsmith_amd_progress", default=None)


class Cancelled(Exception):
    """Raised from a progress callback to abandon the running solve (progress.py:25-26)."""


def report(stage: str, fraction: float | None = None, **info) -> None:
    cb = _callback.get()
    if cb is not None:
        cb({"stage": stage, "fraction": fraction, **info})


@contextmanager
def reporting(callback):
    token = _callback.set(callback)
    try:
        yield
    finally:
        _callback.reset(token)


@dataclass
class RunOptions:
    """The FDTDOptions fields run control reads (options.py:28-92)."""

    cuda_graphs: bool = True
    run_time_fs: float | None = None
    decay_by: float = 1e-5
    max_time_fs: float = 2000.0


class DecayTracker:
    """The per-chunk decisions of run_until_decayed (setup.py:722-773) for one run: divergence, the decay level
    over a sliding window of chunk energies, the stop test, the time-cap warning."""

    def __init__(self, opts, *, max_steps: int, min_steps: int, window: int):
        self.opts = opts
        self.max_steps, self.min_steps = int(max_steps), int(min_steps)
        self.win = deque(maxlen=max(3, int(window)))
        self.peak = self.peak_ref = self.grown_ref = 0.0
        self.level = float("nan")
        self.ref_steps = min(self.min_steps, self.max_steps // 2)

    def update(self, steps: int, e: float, e_ref: float) -> bool:
        """Record the energies after `steps` steps; True when the run is over (decayed or at its cap). Raises
        FloatingPointError on non-finite or runaway energy."""
        finite = np.isfinite(e) and np.isfinite(e_ref)
        if not finite or (steps > self.ref_steps and e > DIVERGED_GROWTH * self.grown_ref > 0):
            why = "non-finite fields" if not finite else \
                f"monitor energy grew to {e / self.grown_ref:.0e} times its peak while the pulse passed"
            raise FloatingPointError(
                f"FDTD diverged after {steps} steps ({why}); the results would be meaningless. "
                "Use a finer mesh (smaller spacing_nm or larger cells_per_wavelength), lower courant, "
                "or give lossy/dispersive materials as Drude/DrudeLorentz")
        self.win.append(e)
        self.peak, self.peak_ref = max(self.peak, e), max(self.peak_ref, e_ref)
        if steps <= self.ref_steps:
            self.grown_ref = self.peak
        floor = max(self.peak, 1e-6 * self.peak_ref)
        self.level = max(self.win) / floor if floor > 0 else float("nan")
        if not self.opts.run_time_fs and steps >= self.min_steps and len(self.win) == self.win.maxlen and \
                self.level < self.opts.decay_by:
            return True
        return steps >= self.max_steps

    def cap_warning(self, steps: int):
        if steps >= self.max_steps and not self.opts.run_time_fs and not (self.level < self.opts.decay_by):
            return (f"fields had decayed only to {self.level:.1e} of peak at the time cap "
                    f"({self.opts.max_time_fs} fs); raise max_time_fs or decay_by")
        return None


def _energies(en) -> tuple[float, float]:
    en = en.detach().cpu().numpy() if hasattr(en, "detach") else np.asarray(en)
    return float(en[0]), float(en[1])


def run_until_decayed(eng, opts, *, max_steps: int, min_steps: int, window: int, warnings: list):
    """Chunks until the energy of the monitors with energy=1 has fallen below opts.decay_by of its peak over a
    window of `window` chunks (after min_steps), or until max_steps / a fixed run time (setup.py:722-773).
    Monitors with energy=2 give a reference: the peak is taken as at least 1e-6 of theirs. Raises
    FloatingPointError on non-finite or runaway energy. Returns (steps, decay_level, used_graphs)."""
    run, graphs = eng.make_runner(opts.cuda_graphs)
    tr = DecayTracker(opts, max_steps=max_steps, min_steps=min_steps, window=window)
    steps = 0
    try:
        while True:
            run()
            steps += eng.stride
            e, e_ref = _energies(eng.energy)  # one sync per chunk
            done = tr.update(steps, e, e_ref)
            report("fdtd_steps", 0.95 * steps / max_steps if opts.run_time_fs else None, steps=steps,
                   max_steps=max_steps)
            if done:
                break
        w = tr.cap_warning(steps)
        if w:
            warnings.append(w)
    finally:
        eng.release()
    return steps, tr.level, graphs


def run_batch_until_decayed(engines, opts, *, max_steps, min_steps, windows, warnings: list, schedule=None,
                            tune: bool = False):
    """run_until_decayed for many runs stepped as one Batch (one DFT stride): each run stops at its own decay
    point or cap, exactly as when run alone -- a finished run leaves the batch (the batch is rebuilt over the
    others), so its monitors hold what the sequential loop (periodic.py:503-519) would give. One readback of every
    member's energy per chunk. Returns [(steps, decay_level, used_graphs)] in the order of `engines`."""
    import torch

    from .batch import Batch

    n = len(engines)
    if schedule is None:
        schedule = filling_tiling([e.shape for e in engines]) + (0,)
    trackers = [DecayTracker(opts, max_steps=max_steps[q], min_steps=min_steps[q], window=windows[q])
                for q in range(n)]
    steps = [0] * n
    graphs = [False] * n
    active = list(range(n))
    stride = engines[0].stride
    while active:
        b = Batch([engines[q] for q in active], schedule=schedule)
        if tune:
            b.autotune()
            schedule = b.schedule
        run, used = b.make_runner(opts.cuda_graphs)
        finished = []
        try:
            while not finished:
                run()
                en = torch.stack([engines[q].energy for q in active]).cpu().numpy()  # one sync per chunk
                for idx, q in enumerate(active):
                    steps[q] += stride
                    graphs[q] = used
                    if trackers[q].update(steps[q], float(en[idx, 0]), float(en[idx, 1])):
                        finished.append(q)
                report("fdtd_steps", None, runs_active=len(active), steps=max(steps))
        finally:
            b.release()
            b.close()
        for q in finished:
            w = trackers[q].cap_warning(steps[q])
            if w:
                warnings.append(f"run {q}: {w}")
        active = [q for q in active if q not in finished]
    return [(steps[q], trackers[q].level, graphs[q]) for q in range(n)]


def flux_points(lo, hi, shape, dx: float, open_lo=()):
    """Closed box with faces on E-node planes lo[d], hi[d] (setup.py:432-468): (e_idx, h_idx1, h_idx2, coeff) in
    the reference's flat indexing, outward flux = 0.5 Re sum coeff E conj(H); per face the two tangential E
    components pair with the face-averaged H of the other, trapezoid weights along node directions."""
    E_i, H1, H2, W = [], [], [], []
    for d in range(3):
        a1, a2 = (d + 1) % 3, (d + 2) % 3
        half1, half2 = np.arange(lo[a1], hi[a1]), np.arange(lo[a2], hi[a2])
        nod1, nod2 = np.arange(lo[a1], hi[a1] + 1), np.arange(lo[a2], hi[a2] + 1)
        w1, w2 = np.ones(len(nod1)), np.ones(len(nod2))
        w1[[0, -1]] = 0.5
        w2[[0, -1]] = 0.5
        faces = ((lo[d], -1.0), (hi[d], 1.0))[1 if d in open_lo else 0:]
        for face, sign in faces:
            for e_c, h_c, i1, i2, wt, tsign in ((a1, a2, half1, nod2, w2, 1.0), (a2, a1, nod1, half2, w1, -1.0)):
                A, B = np.meshgrid(i1, i2, indexing="ij")
                wgt = np.broadcast_to(wt[None, :] if tsign > 0 else wt[:, None], A.shape)
                idx = [None] * 3
                idx[a1], idx[a2] = A.ravel(), B.ravel()
                idx[d] = np.full(A.size, face)
                E_i.append(flat_index(e_c, *idx, shape))
                idx[d] = np.full(A.size, face - 1)
                H1.append(flat_index(h_c, *idx, shape))
                idx[d] = np.full(A.size, face)
                H2.append(flat_index(h_c, *idx, shape))
                W.append(sign * tsign * dx ** 2 * wgt.ravel())
    return np.concatenate(E_i), np.concatenate(H1), np.concatenate(H2), np.concatenate(W)


def flux(e: np.ndarray, h: np.ndarray, coeff: np.ndarray) -> np.ndarray:
    """0.5 Re sum coeff E H*, per frequency row (setup.py:564-566)."""
    return 0.5 * np.real((e * np.conj(h)) @ coeff)


def plane_points(shape, n_cells, k: int):
    """Flat indices (e, h1, h2) for S_z = E_x H_y - E_y H_x on node plane k over the real cells of a periodic
    cell (periodic.py:335-344): term 1 then term 2."""
    i = np.arange(1, n_cells[0] + 1)
    j = np.arange(1, n_cells[1] + 1)
    gi, gj = (a.ravel() for a in np.meshgrid(i, j, indexing="ij"))
    e = np.concatenate([flat_index(0, gi, gj, k, shape), flat_index(1, gi, gj, k, shape)])
    h1 = np.concatenate([flat_index(1, gi, gj, k - 1, shape), flat_index(0, gi, gj, k - 1, shape)])
    h2 = np.concatenate([flat_index(1, gi, gj, k, shape), flat_index(0, gi, gj, k, shape)])
    return e, h1, h2


def orders(e: np.ndarray, h: np.ndarray, n_cells, spacing, kpar=(0.0, 0.0)):
    """Per diffraction order z flux (n_f, nx, ny) of 0.5 Re(E1 H1* - E2 H2*) over the cell (periodic.py:347-363):
    Bloch fields are demodulated by exp(-i k_par . r) first; FFT bin (m, n) holds order k_par + m b1 + n b2."""
    nx, ny = n_cells
    n = nx * ny
    nf = e.shape[0]
    E1, E2 = e[:, :n].reshape(nf, nx, ny), e[:, n:].reshape(nf, nx, ny)
    H1, H2 = h[:, :n].reshape(nf, nx, ny), h[:, n:].reshape(nf, nx, ny)
    if any(kpar):
        i, j = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
        dmod = np.exp(-1j * (kpar[0] * i * spacing[0] + kpar[1] * j * spacing[1]))
        E1, E2, H1, H2 = E1 * dmod, E2 * dmod, H1 * dmod, H2 * dmod
    f = lambda a: np.fft.fft2(a, axes=(1, 2))  # noqa: E731
    out = 0.5 * np.real(f(E1) * np.conj(f(H1)) - f(E2) * np.conj(f(H2))) / n * spacing[0] * spacing[1]
    mx = np.fft.fftfreq(nx, 1 / nx).astype(int)
    my = np.fft.fftfreq(ny, 1 / ny).astype(int)
    return out, mx, my


def max_steps_for(time_fs: float, dt: float, stride: int) -> int:
    """Steps of a run of time_fs, a whole number of DFT chunks (scattering.py:262-263)."""
    return int(math.ceil(time_fs * FS / dt / stride)) * stride
