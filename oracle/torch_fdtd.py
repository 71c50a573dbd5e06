"""Per-step torch restatement of the oracle Engine's update (float64, differentiable): the reference for the
adjoint gradient, by autograd through every step (test infrastructure only).

Restates Engine.update_H / update_E with CPML (engine.py:630-663), the per-step mode-source corrections
(ports.py:362-372) and the running DFT sample (engine.py:564-592) of ``oracle.fdtd.Engine`` -- real fields, no
periodic axes, no dispersive groups -- with the E coefficients of a design box taken from a density rho:
eps = eps_lo + rho (eps_hi - eps_lo) at the E_x, E_y, E_z nodes of each design cell, cb = (dt/dx) / eps. Every
update is out of place so that autograd can differentiate the whole run.
"""

from __future__ import annotations

import numpy as np
import torch

from .fdtd import Engine, _moved

T = torch.float64


def _cat_set(F: torch.Tensor, R, v: torch.Tensor) -> torch.Tensor:
    """F with F[R] = v (out of place)."""
    out = F.clone()
    out[R] = v
    return out


class TorchRun:
    """Forward run of an oracle Engine's configuration in torch. `eng` must be freshly built (fields zero)."""

    def __init__(self, eng: Engine, lo, hi, eps_lo: float, eps_hi: float):
        if eng.complex or any(eng.periodic) or eng.groups or eng.od is not None:
            raise ValueError("TorchRun restates real fields without periodic axes, groups or off-diagonal terms")
        self.eng = eng
        self.lo, self.hi = tuple(lo), tuple(hi)
        self.eps_lo, self.eps_hi = float(eps_lo), float(eps_hi)
        self.cdx = eng.dt / eng.dx if eng.uniform else eng.dt
        self.box = tuple(slice(a, b) for a, b in zip(self.lo, self.hi))

    def cb(self, rho: torch.Tensor):
        """Full-grid E coefficient per component with the design box from rho (each E_c over its stepped box)."""
        eng = self.eng
        out = []
        eps = self.eps_lo + rho * (self.eps_hi - self.eps_lo)
        for c in range(3):
            full = torch.zeros(eng.shape, dtype=T)
            full[eng.RE[c]] = torch.as_tensor(eng.cbE[c], dtype=T)
            out.append(_cat_set(full, self.box, self.cdx / eps))
        return out

    def run(self, rho: torch.Tensor, n_steps: int):
        """(acc_e, acc_h) of the engine's first DFT monitor after n_steps (sampled every eng.stride steps)."""
        eng = self.eng
        cb = self.cb(rho)
        E = torch.zeros((3,) + eng.shape, dtype=T)
        H = torch.zeros((3,) + eng.shape, dtype=T)
        psiE = [torch.zeros(t[5].shape, dtype=T) for t in eng.pml_E]
        psiH = [torch.zeros(t[5].shape, dtype=T) for t in eng.pml_H]
        m = eng.monitors[0]
        omega = torch.as_tensor(np.asarray(m.omega, dtype=float), dtype=T)
        e_idx = torch.as_tensor(np.asarray(m.e_idx))
        acc_e = torch.zeros((len(omega), len(e_idx)), dtype=torch.complex128)
        has_h = m.h_idx1 is not None
        if has_h:
            h1, h2 = torch.as_tensor(np.asarray(m.h_idx1)), torch.as_tensor(np.asarray(m.h_idx2))
            acc_h = torch.zeros((len(omega), len(h1)), dtype=torch.complex128)
        src = [s for s in eng.sources]
        for n in range(n_steps):
            # H update (engine.py:630-643)
            Hn = []
            for c in range(3):
                R = eng.RH[c]
                d1, d2 = (c + 1) % 3, (c + 2) % 3
                a = E[d2][_moved(R, d1, 1)] - E[d2][R]
                b = E[d1][_moved(R, d2, 1)] - E[d1][R]
                if not eng.uniform:
                    a, b = a * eng.inv_d[d1], b * eng.inv_d[d2]
                Hn.append(_cat_set(H[c], R, H[c][R] - eng.cH * (a - b)))
            for q, (c, s_, d, Rs, bb, cp, kp, _) in enumerate(eng.pml_H):
                der = E[s_][_moved(Rs, d, 1)] - E[s_][Rs]
                psiH[q] = torch.as_tensor(bb, dtype=T) * psiH[q] + torch.as_tensor(cp, dtype=T) * der
                inc = psiH[q] + (torch.as_tensor(kp, dtype=T) * der if kp is not None else 0.0)
                Hn[c] = _cat_set(Hn[c], Rs, Hn[c][Rs] + inc)
            H = torch.stack(Hn)
            for s in src:
                if len(s.h_idx):
                    H = H.reshape(-1).index_add(0, torch.as_tensor(s.h_idx),
                                                torch.as_tensor(s.mH @ s.tE[n], dtype=T)).reshape(H.shape)
            # E update (engine.py:645-663)
            En = []
            for c in range(3):
                R = eng.RE[c]
                d1, d2 = (c + 1) % 3, (c + 2) % 3
                a = H[d2][R] - H[d2][_moved(R, d1, -1)]
                b = H[d1][R] - H[d1][_moved(R, d2, -1)]
                if not eng.uniform:
                    a, b = a * eng.inv_d[d1], b * eng.inv_d[d2]
                En.append(_cat_set(E[c], R, E[c][R] + cb[c][R] * (a - b)))
            for q, (c, s_, d, Rs, bb, cp, kp, _) in enumerate(eng.pml_E):
                der = H[s_][Rs] - H[s_][_moved(Rs, d, -1)]
                psiE[q] = torch.as_tensor(bb, dtype=T) * psiE[q] + torch.as_tensor(cp, dtype=T) * der
                inc = psiE[q] + (torch.as_tensor(kp, dtype=T) * der if kp is not None else 0.0)
                En[c] = _cat_set(En[c], Rs, En[c][Rs] + inc)
            E = torch.stack(En)
            for s in src:
                if len(s.e_idx):
                    E = E.reshape(-1).index_add(0, torch.as_tensor(s.e_idx),
                                                torch.as_tensor(s.mE @ s.tH[n], dtype=T)).reshape(E.shape)
            # DFT sample after each chunk (engine.py:564-592)
            if (n + 1) % eng.stride == 0:
                te = (n + 1) * eng.dt
                acc_e = acc_e + torch.exp(1j * omega * te)[:, None] * E.reshape(-1)[e_idx][None, :]
                if has_h:
                    sh = 0.5 * (H.reshape(-1)[h1] + H.reshape(-1)[h2])
                    acc_h = acc_h + torch.exp(1j * omega * (te - 0.5 * eng.dt))[:, None] * sh[None, :]
        return acc_e, (acc_h if has_h else None)
