"""The drop-in at lightsmith's plugin boundary: a ``Backend`` named ``lightsmith-fdtd`` (method ``fdtd``) that solves
scattering problems, layer stacks and periodic structures on the HIP engine (backends/base.py:20-64 and
backends/fdtd.py:55-158 of the reference).

Installed next to lightsmith, the class subclasses the reference's own ``Backend`` and registers through the
``lightsmith.backends`` entry-point group; plugins load after the built-ins and registration keys on ``name``
(base.py:91-107), so it takes the built-in FDTD's place and ``ls.solve(sim, method="fdtd", device="gpu")``
reaches it unchanged. Without lightsmith it subclasses the mirror below (same class attributes and classmethods)
and ``lightsmith_amd.api.solve`` is the front door.

The backend takes lightsmith's objects as they are (a reference ``Simulation``, ``LayerStackProblem`` or
``PeriodicProblem`` is read through its JSON, the schemas of ``lightsmith_amd.problem``) and returns a ``Result`` or a
``SpectrumResult`` with the reference's fields (the reference's own classes when they are importable). It runs on a GPU only: a CPU device is refused, and the HIP library must load.
"""

from __future__ import annotations

from abc import ABC, abstractmethod
from typing import Any, ClassVar

from . import problem as P

try:  # the reference's registry, when lightsmith is installed
    from lightsmith.backends.base import Backend as _RefBackend  # type: ignore
except Exception:  # noqa: BLE001 - lightsmith absent (or unimportable): use the mirror
    _RefBackend = None


class Backend(ABC):
    """Mirror of lightsmith's backend interface (base.py:20-64)."""

    name: ClassVar[str]
    method: ClassVar[str]
    license: ClassVar[str]
    citation: ClassVar[str] = ""
    gpu: ClassVar[bool] = False
    differentiable: ClassVar[bool] = False
    problem_kinds: ClassVar[tuple[str, ...]] = ("scattering",)
    options_model: ClassVar[type | None] = None

    @classmethod
    def available(cls) -> tuple[bool, str]:
        return True, ""

    @classmethod
    @abstractmethod
    def supports(cls, sim) -> tuple[bool, str]:
        ...

    @classmethod
    def estimate(cls, sim, device) -> dict[str, Any]:
        return {}

    @classmethod
    @abstractmethod
    def solve(cls, sim, device):
        ...

    @classmethod
    def info(cls) -> dict[str, Any]:
        ok, why = cls.available()
        return {"name": cls.name, "method": cls.method, "problem_kinds": list(cls.problem_kinds),
                "license": cls.license, "gpu": cls.gpu, "differentiable": cls.differentiable,
                "available": ok, "unavailable_reason": why or None, "citation": cls.citation}


_REGISTRY: dict[str, type] = {}


def register(cls):
    """base.py:91-93: keyed on the backend's name."""
    _REGISTRY[cls.name] = cls
    return cls


def get_backend(name_or_method: str, kind: str | None = None):
    """By backend name, else the first available backend of that method for the problem kind (base.py:110-124)."""
    if name_or_method in _REGISTRY:
        return _REGISTRY[name_or_method]
    cands = [b for b in _REGISTRY.values() if b.method == name_or_method]
    if kind is not None:
        cands = [b for b in cands if kind in b.problem_kinds] or cands
    if not cands:
        raise KeyError(f"unknown method or backend {name_or_method!r}; known: {sorted(_REGISTRY)}")
    for b in cands:
        if b.available()[0]:
            return b
    return cands[0]


def _device_name(device) -> str:
    """The torch device string of a lightsmith Device ("torch", "cuda:0"), a torch.device or a string."""
    if hasattr(device, "framework") and hasattr(device, "name"):
        if device.framework != "torch":
            raise ValueError(f"lightsmith-fdtd (HIP) runs on a torch GPU device, got {device.framework}")
        return str(device.name)
    name = str(device)
    return {"gpu": "cuda:0", "auto": "cuda:0"}.get(name.lower(), name)


def _unsupported(sim) -> str:
    """Why this problem is not on the HIP path ('' if it is): backends/base.py:67-86 and fdtd.py:65-85."""
    missing = []
    out = sim.outputs
    if out.far_field is not None and sim.substrate is not None:
        missing.append("far_field output over a substrate")
    if sim.substrate is not None:
        from .fdtd.scattering import substrate_unsupported
        why = substrate_unsupported(sim)
        if why:
            missing.append(f"this substrate ({why})")
    if sim.orientation_average:
        missing.append("orientation_average")
    if getattr(sim.illumination, "kind", "plane_wave") != "plane_wave":
        missing.append(f"{sim.illumination.kind} illumination")
    return "does not support " + ", ".join(missing) if missing else ""


KINDS = ("scattering", "periodic", "layer_stack", "emission")


def _kind(sim) -> str:
    return sim.get("kind", "scattering") if isinstance(sim, dict) else getattr(sim, "kind", "scattering")


def _as_problem(sim):
    if _kind(sim) not in KINDS:
        raise ValueError(f"lightsmith-fdtd (HIP) solves {', '.join(KINDS)} problems; got {_kind(sim)!r}")
    return P.parse_problem(sim)


def _result(**kw):
    try:
        from lightsmith.results import Result as RefResult  # type: ignore
        return RefResult.from_arrays(**kw)
    except Exception:  # noqa: BLE001 - lightsmith absent
        return P.Result.from_arrays(**kw)


def _emission(res: P.EmissionResult):
    try:
        from lightsmith.results import EmissionResult as RefEmission  # type: ignore
        return RefEmission.model_validate(res.model_dump(mode="json"))
    except Exception:  # noqa: BLE001 - lightsmith absent
        return res


def _spectrum(res: P.SpectrumResult):
    try:
        from lightsmith.results import SpectrumResult as RefSpectrum  # type: ignore
        return RefSpectrum.model_validate(res.model_dump(mode="json"))
    except Exception:  # noqa: BLE001 - lightsmith absent
        return res


class _NativeFDTD:
    """lightsmith's FDTD backend on MI355X (fdtd.py:55-158): scattering problems, layer stacks and periodic
    structures, float32 fields on the GPU."""

    name = "lightsmith-fdtd"
    method = "fdtd"
    license = "MIT"
    citation = ("Taflove & Hagness, Computational Electrodynamics (2005); Roden & Gedney, MOTL 27, 334 (2000); "
                "Oskooi et al., CPC 181, 687 (2010)")
    gpu = True
    differentiable = False
    problem_kinds = KINDS
    options_model = P.FDTDOptions

    @classmethod
    def available(cls) -> tuple[bool, str]:
        try:
            from . import _lib
            _lib.load()
        except Exception as e:  # noqa: BLE001
            return False, f"the HIP library did not load: {e}"
        return True, ""

    @classmethod
    def supports(cls, sim) -> tuple[bool, str]:
        if _kind(sim) not in KINDS:
            return False, "lightsmith-fdtd (HIP) solves scattering, periodic, layer_stack and emission problems"
        try:
            s = _as_problem(sim)
        except Exception as e:  # noqa: BLE001 - a reference feature the mirror schema lacks
            return False, f"problem not on the HIP path: {e}"
        if s.kind == "emission":
            from .fdtd import emission
            from .fdtd.scattering import particle_media
            try:
                cls._options(s)
                ok, why = emission.supports(s)
                if ok:
                    particle_media(s.particles, s.medium_index, s.wavelength_array(), [])
            except ValueError as e:
                return False, str(e)
            return ok, why
        if s.kind != "scattering":
            from .fdtd import spectra
            try:
                cls._options(s)
                ok, why = spectra.supports(s)
                if ok:
                    spectra.MediaTable(s, s.wavelength_array(), [])
            except ValueError as e:
                return False, str(e)
            return ok, why
        why = _unsupported(s)
        if why:
            return False, why
        try:
            cls._options(s)
            from .fdtd.scattering import scene_media
            scene_media(s, [])
        except ValueError as e:
            return False, str(e)
        return True, ""

    @classmethod
    def _options(cls, sim) -> P.FDTDOptions:
        opts = P.FDTDOptions.model_validate(P.options_for(sim, cls.name, cls.method))
        if opts.precision == "double":
            raise ValueError("precision='double' runs on the CPU; the HIP engine steps float32 fields")
        if opts.grid == "auto" or opts.mesh_override:
            raise ValueError("graded grids (grid='auto', mesh_override) are not on the HIP path; use grid='uniform'")
        return opts

    @classmethod
    def _estimate_spectrum(cls, s) -> dict[str, Any]:
        from .fdtd import spectra
        from .fdtd.run import FS
        opts = cls._options(s)
        p = spectra.plan(s, opts)
        c = p["cell"]
        out = {"grid": list(c.shape), "real_cells": list(c.n_cells) + [c.shape[2]], "cells": c.cells,
               "spacing_nm": list(c.spacing), "dt_fs": p["dt"] / FS, "steps_max": p["max_steps"], "runs": p["runs"],
               "dft_stride": p["dft_stride"], "precision": "single", "memory_bytes": p["memory_bytes"],
               "fits": c.cells <= opts.max_cells, "warnings": p["warnings"]}
        if p["oblique_runs"]:
            out["oblique_runs"] = p["oblique_runs"]
            out["notes"] = ("oblique incidence: one fixed-k_par complex-field run per (angle, wavelength, "
                            "polarisation), all stepped together as one batch")
        return out

    @classmethod
    def estimate(cls, sim, device) -> dict[str, Any]:
        from .fdtd import scattering
        s = _as_problem(sim)
        if s.kind == "emission":
            from .fdtd import emission
            p = emission.plan(s, cls._options(s))
            lay = p["layout"]
            return {"grid": list(lay.shape), "cells": lay.cells, "spacing_nm": p["dx"],
                    "dt_fs": p["dt"] / scattering.FS, "steps_estimate": p["est_steps"], "steps_max": p["max_steps"],
                    "runs": p["runs"], "dft_stride": p["stride"], "precision": "single",
                    "memory_bytes": p["memory_bytes"], "fits": lay.cells <= cls._options(s).max_cells,
                    "warnings": p["warnings"],
                    "notes": "two runs, stepped as one batch: the scene, and the same dipole in the medium alone (P0)"}
        if s.kind != "scattering":
            return cls._estimate_spectrum(s)
        opts = cls._options(s)
        p = scattering.plan(s, opts)
        lay = p["layout"]
        return {"grid": list(lay.shape), "cells": lay.cells, "spacing_nm": p["dx"],
                "dt_fs": p["dt"] / scattering.FS, "steps_estimate": p["est_steps"], "steps_max": p["max_steps"],
                "dft_stride": p["stride"], "precision": "single", "memory_bytes": p["memory_bytes"],
                "fits": lay.cells <= opts.max_cells, "warnings": p["warnings"], "symmetry": "off"}

    @classmethod
    def solve(cls, sim, device):
        ok, why = cls.supports(sim)
        if not ok:
            raise ValueError(f"{cls.name} {why}")
        from .fdtd import scattering
        s = _as_problem(sim)
        name = _device_name(device)
        if not name.startswith("cuda"):
            raise ValueError(f"{cls.name} runs on a GPU (device 'gpu' or 'cuda:N'), got {name!r}")
        if s.kind == "emission":
            from .fdtd import emission
            return _emission(emission.solve(s, cls._options(s), name))
        if s.kind != "scattering":
            from .fdtd import spectra
            return _spectrum(spectra.solve(s, cls._options(s), name))
        out = scattering.solve(s, cls._options(s), name)
        res = _result(method=cls.method, backend=cls.name, wavelengths_nm=s.wavelength_array(), c_ext=out["c_ext"],
                      c_sca=out["c_sca"], c_abs=out["c_abs"], r_eq=s.equivalent_radius_nm(),
                      provenance=out["provenance"], warnings=out["warnings"])
        for key, ref_cls in (("far_field", "FarFieldData"), ("near_field", "NearFieldData"),
                             ("resonances", "ResonanceData")):
            if out[key] is not None:
                data = out[key]
                if not isinstance(res, P.Result):  # the reference's Result: its own data classes
                    import lightsmith.results as ref_results  # type: ignore
                    data = [getattr(ref_results, ref_cls).model_validate(d.model_dump()) for d in data]
                setattr(res, key, data)
        return res


_Base = _RefBackend if _RefBackend is not None else Backend


@register
class NativeFDTD(_NativeFDTD, _Base):
    """lightsmith's FDTD backend on MI355X (fdtd.py:55-158): scattering problems, layer stacks and periodic
    structures, float32 fields on the GPU."""


__all__ = ["Backend", "NativeFDTD", "get_backend", "register"]
