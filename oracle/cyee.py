"""ctypes front end of oracle/cyee.c: an oracle Engine's configuration stepped in plain C (TEST INFRASTRUCTURE ONLY).

``CYee(eng)`` takes a freshly configured ``oracle.fdtd.Engine`` (float64, real fields, uniform grid, no periodic
axes, pure D-form groups, point dipoles, DFT monitors on the grid) and copies its state into arrays of its own;
``run(n, stride)`` steps them in C (OpenMP over the slowest axis) in the oracle's operation order. It serves the
multi-core CPU baseline and checks at sizes where NumPy is slow; tests/test_oracle_pins.py pins it to the NumPy
oracle (fields equal to 1e-13, DFT spectra to 1e-12).

The library is built by ``build()`` of the repository's entry point (``oracle/Makefile``: gcc -O2 -fopenmp
-ffp-contract=off) into ``oracle/_cyee.so``.
"""

from __future__ import annotations

import ctypes as C
import os

import numpy as np

from .fdtd import Engine, PureGroup
from .sources import PointDipole

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_cyee.so")
MAX_POLES = 8

_i64p = C.POINTER(C.c_longlong)
_dp = C.POINTER(C.c_double)


class _Box(C.Structure):
    _fields_ = [("lo", C.c_int * 3), ("hi", C.c_int * 3)]


class _Pml(C.Structure):
    _fields_ = [("c", C.c_int), ("src", C.c_int), ("d", C.c_int), ("r", _Box), ("b", _dp), ("cp", _dp), ("kp", _dp),
                ("psi", _dp)]


class _Group(C.Structure):
    _fields_ = [("n", C.c_longlong), ("idx", _i64p), ("inv_eps_inf", C.c_double), ("n_poles", C.c_int),
                ("coef", (C.c_double * 3) * MAX_POLES), ("D", _dp), ("E_prev", _dp), ("P", _dp)]


class _Point(C.Structure):
    _fields_ = [("n", C.c_longlong), ("idx", _i64p), ("coef", _dp), ("wave", _dp), ("n_wave", C.c_longlong)]


class _Monitor(C.Structure):
    _fields_ = [("n_freq", C.c_int), ("n_e", C.c_longlong), ("n_h", C.c_longlong), ("omega", _dp), ("e_idx", _i64p),
                ("h1", _i64p), ("h2", _i64p), ("acc_e", _dp), ("acc_h", _dp), ("energy", C.c_int)]


class _Engine(C.Structure):
    _fields_ = [("n", C.c_int * 3), ("E", _dp), ("H", _dp), ("rE", _Box * 3), ("rH", _Box * 3), ("cb", _dp * 3),
                ("cH", C.c_double), ("dt", C.c_double), ("n_pml_h", C.c_int), ("n_pml_e", C.c_int),
                ("n_groups", C.c_int), ("n_points", C.c_int), ("n_monitors", C.c_int), ("pml_h", C.POINTER(_Pml)),
                ("pml_e", C.POINTER(_Pml)), ("groups", C.POINTER(_Group)), ("points", C.POINTER(_Point)),
                ("monitors", C.POINTER(_Monitor)), ("energy", C.c_double * 2), ("step", C.c_longlong)]


_lib = None


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise RuntimeError(f"{LIB} is not built (build() of the repository entry point, or make -C oracle)")
        _lib = C.CDLL(LIB)
        _lib.cy_run.argtypes = [C.POINTER(_Engine), C.c_longlong, C.c_int]
        _lib.cy_run.restype = None
        _lib.cy_abi_version.restype = C.c_int
    return _lib


def _box(R) -> _Box:
    b = _Box()
    for a, s in enumerate(R):
        b.lo[a], b.hi[a] = s.start, s.stop
    return b


class CYee:
    """State copied from an oracle Engine and stepped in C."""

    def __init__(self, eng: Engine):
        if eng.rdt != np.float64 or eng.complex or any(eng.periodic) or not eng.uniform or eng.od is not None:
            raise ValueError("CYee restates float64 real fields on a uniform grid without periodic axes or "
                             "off-diagonal rows")
        if eng.time_monitors:
            raise ValueError("CYee has no time monitors")
        self.lib = load()
        keep = self._keep = []

        def arr(a, dtype=np.float64):
            a = np.ascontiguousarray(np.array(a, dtype=dtype, copy=True))
            keep.append(a)
            return a

        def dp(a):
            return a.ctypes.data_as(_dp)

        def ip(a):
            return a.ctypes.data_as(_i64p)

        self.shape = eng.shape
        self.E, self.H = arr(eng.E), arr(eng.H)
        e = self._e = _Engine()
        e.n[:] = eng.shape
        e.E, e.H = dp(self.E), dp(self.H)
        for c in range(3):
            e.rE[c], e.rH[c] = _box(eng.RE[c]), _box(eng.RH[c])
            cb = np.broadcast_to(np.asarray(eng.cbE[c], np.float64),
                                 tuple(s.stop - s.start for s in eng.RE[c]))
            e.cb[c] = dp(arr(cb))
        e.cH, e.dt = float(eng.cH), float(eng.dt)
        self.psi_h, self.psi_e = [], []
        for name, terms, store in (("h", eng.pml_H, self.psi_h), ("e", eng.pml_E, self.psi_e)):
            ts = (_Pml * max(1, len(terms)))()
            for q, (c, src, d, Rs, b, cp, kp, psi) in enumerate(terms):
                shape = tuple(s.stop - s.start for s in Rs)
                t = ts[q]
                t.c, t.src, t.d, t.r = c, src, d, _box(Rs)
                t.b = dp(arr(np.asarray(b).reshape(-1)))
                t.cp = dp(arr(np.broadcast_to(cp, shape)))
                t.kp = dp(arr(np.broadcast_to(kp, shape))) if kp is not None else None
                p = arr(np.broadcast_to(psi, shape))
                store.append(p)
                t.psi = dp(p)
            keep.append(ts)
            setattr(e, f"n_pml_{name}", len(terms))
            setattr(e, f"pml_{name}", ts)
        gs = (_Group * max(1, len(eng.groups)))()
        for q, g in enumerate(eng.groups):
            if not isinstance(g, PureGroup):
                raise ValueError("CYee steps pure D-form groups only")
            if len(g.coef) > MAX_POLES:
                raise ValueError(f"at most {MAX_POLES} poles")
            G = gs[q]
            G.n = len(g.idx)
            G.idx = ip(arr(g.idx, np.int64))
            G.inv_eps_inf = float(g.inv_eps_inf)
            G.n_poles = len(g.coef)
            for s, cf in enumerate(g.coef):
                G.coef[s][:] = [float(v) for v in cf]
            G.D, G.E_prev = dp(arr(g.D)), dp(arr(g.E_prev))
            P = arr(np.stack([np.stack(pair) for pair in g.P]) if g.P else np.zeros(0))
            G.P = dp(P)
        keep.append(gs)
        e.n_groups, e.groups = len(eng.groups), gs
        ps = (_Point * max(1, len(eng.sources)))()
        for q, s in enumerate(eng.sources):
            if not isinstance(s, PointDipole):
                raise ValueError("CYee injects point dipoles only")
            P = ps[q]
            P.n = len(s.idx)
            P.idx = ip(arr(s.idx, np.int64))
            P.coef = dp(arr(s._coef))
            w = arr(s._wave)
            P.wave, P.n_wave = dp(w), len(w)
        keep.append(ps)
        e.n_points, e.points = len(eng.sources), ps
        ms = (_Monitor * max(1, len(eng.monitors)))()
        self.acc = []
        for q, m in enumerate(eng.monitors):
            if m.source is not None:
                raise ValueError("CYee samples monitors on the grid only")
            M = ms[q]
            om = arr(m.omega)
            M.n_freq, M.omega = len(om), dp(om)
            M.n_e, M.e_idx = len(m.e_idx), ip(arr(m.e_idx, np.int64))
            ae = arr(np.asarray(m.acc_e).view(np.float64).reshape(len(om), -1))
            M.acc_e = dp(ae)
            ah = None
            if m.acc_h is not None:
                M.n_h, M.h1, M.h2 = len(m.h_idx1), ip(arr(m.h_idx1, np.int64)), ip(arr(m.h_idx2, np.int64))
                ah = arr(np.asarray(m.acc_h).view(np.float64).reshape(len(om), -1))
                M.acc_h = dp(ah)
                M.energy = int(m.energy or 0)
            self.acc.append((ae, ah))
        keep.append(ms)
        e.n_monitors, e.monitors = len(eng.monitors), ms
        e.energy[:] = [float(v) for v in eng.energy]
        e.step = int(eng.n)
        self.stride = int(eng.stride)

    @property
    def n(self) -> int:
        return int(self._e.step)

    @property
    def energy(self) -> np.ndarray:
        return np.array(self._e.energy[:])

    def run(self, n_steps: int, sample: bool = True):
        """n_steps steps; a DFT sample after every multiple of the engine's stride when `sample`."""
        self.lib.cy_run(C.byref(self._e), int(n_steps), self.stride if sample else 0)

    def results(self, q: int = 0):
        ae, ah = self.acc[q]
        return ae.view(np.complex128), (ah.view(np.complex128) if ah is not None else None)
