"""Per-step torch restatement of the oracle Engine's update (float64, differentiable): the reference for the
adjoint gradient, by autograd through every step (test infrastructure only).

Restates Engine.update_H / update_E with CPML (engine.py:630-663), the per-step mode-source corrections
(ports.py:362-372), pure D-form (ADE) groups (engine.py:759-793) and the running DFT sample (engine.py:564-592) of
``oracle.fdtd.Engine`` -- real fields, no periodic axes, no mixed or vertex-tensor groups -- with the E coefficients of a design box taken from a density rho:
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

    def __init__(self, eng: Engine, lo, hi, eps_lo: float, eps_hi: float, device="cpu"):
        from .fdtd import PureGroup
        if eng.complex or any(eng.periodic) or eng.od is not None or \
                any(not isinstance(g, PureGroup) for g in eng.groups):
            raise ValueError("TorchRun restates real fields without periodic axes, mixed/tensor groups or "
                             "off-diagonal terms")
        self.eng = eng
        self.dev = torch.device(device)
        self.lo, self.hi = tuple(lo), tuple(hi)
        self.eps_lo, self.eps_hi = float(eps_lo), float(eps_hi)
        self.cdx = eng.dt / eng.dx if eng.uniform else eng.dt
        self.box = tuple(slice(a, b) for a, b in zip(self.lo, self.hi))
        d = self.dev
        conv = lambda a: None if a is None else torch.as_tensor(a, dtype=T, device=d)  # noqa: E731
        self.pml_H = [(c, s_, dd, Rs, conv(bb), conv(cp), conv(kp)) for c, s_, dd, Rs, bb, cp, kp, _ in eng.pml_H]
        self.pml_E = [(c, s_, dd, Rs, conv(bb), conv(cp), conv(kp)) for c, s_, dd, Rs, bb, cp, kp, _ in eng.pml_E]
        self.groups = [(torch.as_tensor(g.idx, device=d), float(g.inv_eps_inf), [tuple(float(v) for v in c)
                        for c in g.coef]) for g in eng.groups]
        self.src = []
        for s in eng.sources:
            self.src.append((torch.as_tensor(np.asarray(s.h_idx, dtype=np.int64), device=d), conv(s.mH @ s.tE.T).T
                             if len(s.h_idx) else None,
                             torch.as_tensor(np.asarray(s.e_idx, dtype=np.int64), device=d), conv(s.mE @ s.tH.T).T
                             if len(s.e_idx) else None))

    def cb(self, rho: torch.Tensor):
        """Full-grid E coefficient per component with the design box from rho (each E_c over its stepped box)."""
        eng = self.eng
        out = []
        eps = self.eps_lo + rho * (self.eps_hi - self.eps_lo)
        for c in range(3):
            full = torch.zeros(eng.shape, dtype=T, device=self.dev)
            full[eng.RE[c]] = torch.as_tensor(eng.cbE[c], dtype=T, device=self.dev)
            out.append(_cat_set(full, self.box, self.cdx / eps))
        return out

    def run(self, rho: torch.Tensor, n_steps: int):
        """(acc_e, acc_h) of the engine's first DFT monitor after n_steps (sampled every eng.stride steps).
        Differentiable in rho; under torch.no_grad() the updates run in place."""
        eng, d = self.eng, self.dev
        inplace = not torch.is_grad_enabled()
        put = (lambda F, R, v: F.__setitem__(R, v) or F) if inplace else _cat_set  # noqa: E731
        cb = self.cb(rho.to(d))
        E = torch.zeros((3,) + eng.shape, dtype=T, device=d)
        H = torch.zeros((3,) + eng.shape, dtype=T, device=d)
        psiE = [torch.zeros(t[5].shape, dtype=T, device=d) for t in self.pml_E]
        psiH = [torch.zeros(t[5].shape, dtype=T, device=d) for t in self.pml_H]
        gstate = [(torch.zeros(len(idx), dtype=T, device=d), torch.zeros(len(idx), dtype=T, device=d),
                   [(torch.zeros(len(idx), dtype=T, device=d), torch.zeros(len(idx), dtype=T, device=d))
                    for _ in coef]) for idx, _, coef in self.groups]
        m = eng.monitors[0]
        omega = torch.as_tensor(np.asarray(m.omega, dtype=float), dtype=T, device=d)
        e_idx = torch.as_tensor(np.asarray(m.e_idx), device=d)
        acc_e = torch.zeros((len(omega), len(e_idx)), dtype=torch.complex128, device=d)
        has_h = m.h_idx1 is not None
        if has_h:
            h1 = torch.as_tensor(np.asarray(m.h_idx1), device=d)
            h2 = torch.as_tensor(np.asarray(m.h_idx2), device=d)
            acc_h = torch.zeros((len(omega), len(h1)), dtype=torch.complex128, device=d)
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
                Hn.append(put(H[c], R, H[c][R] - eng.cH * (a - b)))
            for q, (c, s_, dd, Rs, bb, cp, kp) in enumerate(self.pml_H):
                der = E[s_][_moved(Rs, dd, 1)] - E[s_][Rs]
                psiH[q] = bb * psiH[q] + cp * der
                inc = psiH[q] + (kp * der if kp is not None else 0.0)
                Hn[c] = put(Hn[c], Rs, Hn[c][Rs] + inc)
            H = H if inplace else torch.stack(Hn)
            for h_idx, mh, _, _ in self.src:
                if mh is not None:
                    H = H.reshape(-1).index_add(0, h_idx, mh[n]).reshape(H.shape)
            # E update (engine.py:645-663)
            En = []
            for c in range(3):
                R = eng.RE[c]
                d1, d2 = (c + 1) % 3, (c + 2) % 3
                a = H[d2][R] - H[d2][_moved(R, d1, -1)]
                b = H[d1][R] - H[d1][_moved(R, d2, -1)]
                if not eng.uniform:
                    a, b = a * eng.inv_d[d1], b * eng.inv_d[d2]
                En.append(put(E[c], R, E[c][R] + cb[c][R] * (a - b)))
            for q, (c, s_, dd, Rs, bb, cp, kp) in enumerate(self.pml_E):
                der = H[s_][Rs] - H[s_][_moved(Rs, dd, -1)]
                psiE[q] = bb * psiE[q] + cp * der
                inc = psiE[q] + (kp * der if kp is not None else 0.0)
                En[c] = put(En[c], Rs, En[c][Rs] + inc)
            E = E if inplace else torch.stack(En)
            for _, _, e_src, me in self.src:
                if me is not None:
                    E = E.reshape(-1).index_add(0, e_src, me[n]).reshape(E.shape)
            # pure D-form groups: D += E_now - E_prev; poles from E_prev; E = (D - sum P) / eps_inf
            for q, (idx, inv, coef) in enumerate(self.groups):
                D, E_prev, P = gstate[q]
                D = D + E.reshape(-1)[idx] - E_prev
                tot = 0.0
                newP = []
                for (P0, P1), (c1, c2, c3) in zip(P, coef):
                    Pn = c1 * P0 + c2 * P1 + c3 * E_prev
                    newP.append((Pn, P0))
                    tot = tot + Pn
                e_new = (D - tot) * inv
                E = E.reshape(-1).index_copy(0, idx, e_new).reshape(E.shape)
                gstate[q] = (D, e_new, newP)
            # DFT sample after each chunk (engine.py:564-592)
            if (n + 1) % eng.stride == 0:
                te = (n + 1) * eng.dt
                acc_e = acc_e + torch.exp(1j * omega * te)[:, None] * E.reshape(-1)[e_idx][None, :]
                if has_h:
                    sh = 0.5 * (H.reshape(-1)[h1] + H.reshape(-1)[h2])
                    acc_h = acc_h + torch.exp(1j * omega * (te - 0.5 * eng.dt))[:, None] * sh[None, :]
        return acc_e, (acc_h if has_h else None)
