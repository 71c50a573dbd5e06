"""Front door with lightsmith's signature: ``solve(problem, method="fdtd", device="gpu")`` and ``estimate``
(api.py:143-186 of the reference), over the backends registered in ``lightsmith_amd.backend``.

    from lightsmith_amd import api
    res = api.solve(sim, method="fdtd", device="gpu")   # a Result: wavelengths_nm, c_ext_nm2, c_sca_nm2, ...

``problem`` is a ``lightsmith_amd.problem.Simulation``, a reference ``Simulation``, a dict or its JSON. Progress
events go to ``progress`` as in the reference (stage, fraction, ...); it may raise ``Cancelled``.
"""

from __future__ import annotations

import platform

import numpy as np

from . import __version__
from .backend import get_backend
from .fdtd.run import report, reporting
from .problem import parse_problem


def _resolve(problem, method: str):
    if method == "auto":
        method = "fdtd"
    b = get_backend(method, kind=getattr(problem, "kind", "scattering"))
    for ok, why in (b.available(), b.supports(problem)):
        if not ok:
            raise ValueError(f"backend {b.name!r} cannot run this problem: {why}")
    return b


def solve(problem, method: str = "fdtd", device: str = "gpu", progress=None):
    """Solve ``problem`` with the FDTD backend on the GPU (api.py:143-162)."""
    problem = parse_problem(problem)
    b = _resolve(problem, method)
    with reporting(progress):
        report("solve_started", 0.0, backend=b.name, method=b.method, device=str(device))
        res = b.solve(problem, device)
        report("solve_done", 1.0, backend=b.name)
    import torch
    res.provenance.setdefault("versions", {"lightsmith_amd": __version__, "numpy": np.__version__,
                                           "torch": torch.__version__, "python": platform.python_version()})
    return res


def estimate(problem, method: str = "fdtd", device: str = "gpu") -> dict:
    """Cheap pre-flight estimate (api.py:165-183): grid, cells, dt, steps, stride, memory. Never solves."""
    problem = parse_problem(problem)
    b = get_backend("fdtd" if method == "auto" else method, kind=getattr(problem, "kind", "scattering"))
    return {"backend": b.name, "method": b.method, **b.estimate(problem, device)}


__all__ = ["estimate", "solve"]
