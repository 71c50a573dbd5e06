"""Problem and result types of the FDTD front door: the subset of lightsmith's typed specification that plane-wave
scattering by particles, layer stacks and periodic structures use (spec.py, problems.py, materials.py, results.py,
backends/_fdtd/options.py of the reference), with the same field names, defaults and validation, so that a
reference problem serialised to JSON parses here and a ``Result`` / ``SpectrumResult`` produced here has the
reference's fields.

Units: nm (vacuum wavelengths), exp(-i omega t), cross sections in nm^2, energies in eV.
"""

from __future__ import annotations

from typing import Annotated, Any, Literal, Union

import numpy as np
from pydantic import BaseModel, ConfigDict, Field, field_validator, model_validator

HC_EV_NM = 1239.841984  # h c in eV nm (materials.py:32)


class _Model(BaseModel):
    model_config = ConfigDict(extra="forbid", frozen=True)


# --- materials (materials.py:49-113) -----------------------------------------------------------------------------

class ConstantIndex(_Model):
    kind: Literal["constant_index"] = "constant_index"
    n_real: float
    n_imag: float = Field(0.0, ge=0)

    def eps(self, wavelength_nm):
        wl = np.asarray(wavelength_nm, dtype=float)
        return np.full(wl.shape, complex(self.n_real, self.n_imag) ** 2)


class ConstantPermittivity(_Model):
    kind: Literal["constant_permittivity"] = "constant_permittivity"
    eps_real: float
    eps_imag: float = Field(0.0, ge=0)

    def eps(self, wavelength_nm):
        wl = np.asarray(wavelength_nm, dtype=float)
        return np.full(wl.shape, complex(self.eps_real, self.eps_imag))


class Drude(_Model):
    """eps(w) = eps_inf - wp^2 / (w^2 + i gamma w), energies in eV."""

    kind: Literal["drude"] = "drude"
    eps_inf: float = 1.0
    wp_ev: float = Field(gt=0)
    gamma_ev: float = Field(ge=0)

    def eps(self, wavelength_nm):
        w = HC_EV_NM / np.asarray(wavelength_nm, dtype=float)
        return self.eps_inf - self.wp_ev ** 2 / (w ** 2 + 1j * self.gamma_ev * w)


class LorentzPole(_Model):
    delta_eps: float
    omega0_ev: float = Field(gt=0)
    gamma_ev: float = Field(ge=0)


class DrudeLorentz(_Model):
    """eps(w) = eps_inf - wp^2/(w^2 + i g w) + sum_j d_j w0_j^2 / (w0_j^2 - w^2 - i G_j w)."""

    kind: Literal["drude_lorentz"] = "drude_lorentz"
    eps_inf: float = 1.0
    wp_ev: float = Field(0.0, ge=0)
    gamma_ev: float = Field(0.0, ge=0)
    poles: list[LorentzPole] = Field(default_factory=list)

    def eps(self, wavelength_nm):
        w = HC_EV_NM / np.asarray(wavelength_nm, dtype=float)
        e = self.eps_inf + 0j * w
        if self.wp_ev > 0:
            e = e - self.wp_ev ** 2 / (w ** 2 + 1j * self.gamma_ev * w)
        for p in self.poles:
            e = e + p.delta_eps * p.omega0_ev ** 2 / (p.omega0_ev ** 2 - w ** 2 - 1j * p.gamma_ev * w)
        return e


Material = Annotated[Union[ConstantIndex, ConstantPermittivity, Drude, DrudeLorentz], Field(discriminator="kind")]


# --- shapes (spec.py:29-121) -------------------------------------------------------------------------------------

class Sphere(_Model):
    kind: Literal["sphere"] = "sphere"
    radius_nm: float = Field(gt=0)

    def volume(self) -> float:
        return 4 / 3 * np.pi * self.radius_nm ** 3

    def bounding_half_extent(self) -> np.ndarray:
        return np.full(3, self.radius_nm)


class Spheroid(_Model):
    kind: Literal["spheroid"] = "spheroid"
    semi_axis_xy_nm: float = Field(gt=0)
    semi_axis_z_nm: float = Field(gt=0)

    def volume(self) -> float:
        return 4 / 3 * np.pi * self.semi_axis_xy_nm ** 2 * self.semi_axis_z_nm

    def bounding_half_extent(self) -> np.ndarray:
        return np.array([self.semi_axis_xy_nm, self.semi_axis_xy_nm, self.semi_axis_z_nm])


class Cylinder(_Model):
    kind: Literal["cylinder"] = "cylinder"
    radius_nm: float = Field(gt=0)
    length_nm: float = Field(gt=0)

    def volume(self) -> float:
        return np.pi * self.radius_nm ** 2 * self.length_nm

    def bounding_half_extent(self) -> np.ndarray:
        return np.array([self.radius_nm, self.radius_nm, self.length_nm / 2])


class Shell(_Model):
    thickness_nm: float = Field(gt=0)
    material: Material


class LayeredSphere(_Model):
    kind: Literal["layered_sphere"] = "layered_sphere"
    core_radius_nm: float = Field(gt=0)
    shells: list[Shell] = Field(min_length=1)

    @property
    def radii_nm(self) -> list[float]:
        r = [self.core_radius_nm]
        for s in self.shells:
            r.append(r[-1] + s.thickness_nm)
        return r

    @property
    def radius_nm(self) -> float:
        return self.radii_nm[-1]

    def volume(self) -> float:
        return 4 / 3 * np.pi * self.radius_nm ** 3

    def bounding_half_extent(self) -> np.ndarray:
        return np.full(3, self.radius_nm)


Shape = Annotated[Union[Sphere, Spheroid, Cylinder, LayeredSphere], Field(discriminator="kind")]


class Particle(_Model):
    shape: Shape
    material: Material
    position_nm: tuple[float, float, float] = (0.0, 0.0, 0.0)

    def region_materials(self) -> list:
        """Materials by region: [core/body, shell 1, shell 2, ...]."""
        return [self.material, *[s.material for s in getattr(self.shape, "shells", [])]]


# --- illumination, wavelengths, outputs (spec.py:135-246) ----------------------------------------------------------

class PlaneWave(_Model):
    kind: Literal["plane_wave"] = "plane_wave"
    direction: tuple[float, float, float] = (0.0, 0.0, 1.0)
    polarization: tuple[float, float, float] = (1.0, 0.0, 0.0)

    @model_validator(mode="after")
    def _orthogonal(self):
        k, e = np.asarray(self.direction), np.asarray(self.polarization)
        if np.linalg.norm(k) == 0 or np.linalg.norm(e) == 0:
            raise ValueError("direction and polarization must be non-zero")
        if abs(k @ e) > 1e-9 * np.linalg.norm(k) * np.linalg.norm(e):
            raise ValueError("polarization must be orthogonal to direction")
        return self

    def unit_vectors(self) -> tuple[np.ndarray, np.ndarray]:
        k, e = np.asarray(self.direction, float), np.asarray(self.polarization, float)
        return k / np.linalg.norm(k), e / np.linalg.norm(e)


class WavelengthRange(_Model):
    start_nm: float = Field(gt=0)
    stop_nm: float = Field(gt=0)
    num: int = Field(ge=1)

    def values(self) -> np.ndarray:
        return np.linspace(self.start_nm, self.stop_nm, self.num)


class NearFieldRequest(_Model):
    """The total electric field on a square plane through the scene (spec.py:206-213)."""

    plane: Literal["xy", "xz", "yz"] = "xz"
    offset_nm: float = 0.0  # the plane's position along its normal
    half_width_nm: float = Field(gt=0)
    points: int = Field(101, ge=2)  # samples per axis
    wavelengths_nm: list[float] = Field(min_length=1)


class FarFieldRequest(_Model):
    """dC_sca/dOmega on a theta x phi grid: theta from the incident direction, phi the azimuth around it from the
    incident polarisation (phi = 0 the E-plane, 90 the H-plane) (spec.py:216-225)."""

    theta_deg: list[float] = Field(min_length=1)
    phi_deg: list[float] = Field([0.0, 90.0], min_length=1)
    wavelengths_nm: list[float] = Field(min_length=1)


class ResonanceRequest(_Model):
    """Resonances of the ringing field after the incident pulse has passed: wavelength, Q and relative amplitude of
    the modes in the band (spec.py:228-233)."""

    max_modes: int = Field(10, ge=1, le=100)


class Outputs(_Model):
    """Cross sections, the near and far fields and the resonances (spec.py:236-241)."""

    cross_sections: bool = True
    near_field: NearFieldRequest | None = None
    far_field: FarFieldRequest | None = None
    resonances: ResonanceRequest | None = None


class Substrate(_Model):
    """A semi-infinite half-space z < interface_z_nm of the given material (spec.py:142-146)."""

    material: Material
    interface_z_nm: float = 0.0


class Simulation(_Model):
    """A plane-wave scattering problem (spec.py:251-296)."""

    kind: Literal["scattering"] = "scattering"
    schema_version: Literal["0.1"] = "0.1"
    particles: list[Particle] = Field(min_length=1)
    medium_index: float = Field(1.0, gt=0)
    illumination: PlaneWave = PlaneWave()
    wavelengths: list[float] | WavelengthRange
    substrate: Substrate | None = None
    orientation_average: bool = False
    outputs: Outputs = Outputs()
    options: dict[str, dict[str, Any]] = Field(default_factory=dict)

    @field_validator("wavelengths")
    @classmethod
    def _positive(cls, v):
        if isinstance(v, list) and (not v or min(v) <= 0):
            raise ValueError("wavelengths must be a non-empty list of positive values")
        return v

    def options_for(self, backend_name: str, method: str) -> dict[str, Any]:
        return {**self.options.get(method, {}), **self.options.get(backend_name, {})}

    def wavelength_array(self) -> np.ndarray:
        w = self.wavelengths
        return w.values() if isinstance(w, WavelengthRange) else np.asarray(w, dtype=float)

    def equivalent_radius_nm(self) -> float:
        v = sum(p.shape.volume() for p in self.particles)
        return (3 * v / (4 * np.pi)) ** (1 / 3)


Wavelengths = Union[list[float], WavelengthRange]


def wavelength_array(w) -> np.ndarray:
    """Vacuum wavelengths (nm) of a list or a WavelengthRange (problems.py:34-35)."""
    return w.values() if isinstance(w, WavelengthRange) else np.asarray(w, dtype=float)


# --- 2D profiles, layer stacks and periodic structures (problems.py:41-160) ------------------------------------------

class Rectangle(_Model):
    kind: Literal["rectangle"] = "rectangle"
    center_nm: tuple[float, float] = (0.0, 0.0)
    size_nm: tuple[float, float]


class Circle(_Model):
    kind: Literal["circle"] = "circle"
    center_nm: tuple[float, float] = (0.0, 0.0)
    radius_nm: float = Field(gt=0)


class Ellipse(_Model):
    kind: Literal["ellipse"] = "ellipse"
    center_nm: tuple[float, float] = (0.0, 0.0)
    semi_axes_nm: tuple[float, float]
    angle_deg: float = 0.0


class Polygon(_Model):
    kind: Literal["polygon"] = "polygon"
    vertices_nm: list[tuple[float, float]] = Field(min_length=3)


Profile = Annotated[Union[Rectangle, Circle, Ellipse, Polygon], Field(discriminator="kind")]


class Structure2D(_Model):
    name: str | None = None
    profile: Profile
    material: Material


class Layer(_Model):
    thickness_nm: float = Field(ge=0)
    material: Material
    coherent: bool = True


class LayerStackProblem(_Model):
    """Planar multilayer: reflection, transmission and absorption against wavelength and angle. Layers are listed
    from the superstrate side; light comes in from the superstrate and travels towards the substrate (+z)."""

    kind: Literal["layer_stack"] = "layer_stack"
    superstrate: Material
    layers: list[Layer] = Field(default_factory=list)
    substrate: Material
    wavelengths: Wavelengths
    angles_deg: list[float] = [0.0]
    polarizations: list[Literal["s", "p"]] = ["s", "p"]
    options: dict[str, dict[str, Any]] = Field(default_factory=dict)

    def wavelength_array(self) -> np.ndarray:
        return wavelength_array(self.wavelengths)


class PatternedLayer(_Model):
    thickness_nm: float = Field(gt=0)
    background: Material
    shapes: list[Structure2D] = Field(default_factory=list)


class Incidence(_Model):
    theta_deg: float = 0.0
    phi_deg: float = 0.0
    polarizations: list[Literal["s", "p"]] = ["s", "p"]


class PeriodicProblem(_Model):
    """A 1D or 2D periodic stack of patterned layers under plane-wave illumination (lattice_a2_nm None: a 1D
    grating, invariant along y)."""

    kind: Literal["periodic"] = "periodic"
    lattice_a1_nm: tuple[float, float]
    lattice_a2_nm: tuple[float, float] | None = None
    superstrate: Material
    layers: list[PatternedLayer] = Field(min_length=1)
    substrate: Material
    wavelengths: Wavelengths
    incidence: Incidence = Incidence()
    options: dict[str, dict[str, Any]] = Field(default_factory=dict)

    def wavelength_array(self) -> np.ndarray:
        return wavelength_array(self.wavelengths)


class Dipole(_Model):
    """A point electric dipole emitter (problems.py:269-283)."""

    position_nm: tuple[float, float, float] = (0.0, 0.0, 0.0)
    orientation: tuple[float, float, float] = (0.0, 0.0, 1.0)

    @model_validator(mode="after")
    def _nonzero(self):
        if not np.any(self.orientation):
            raise ValueError("dipole orientation must be non-zero")
        return self

    def unit(self) -> np.ndarray:
        o = np.asarray(self.orientation, dtype=float)
        return o / np.linalg.norm(o)


class RadiationPatternRequest(_Model):
    """Radiated power per solid angle; theta from +z, phi from +x (problems.py:286-291)."""

    theta_deg: list[float] = Field(min_length=1)
    phi_deg: list[float] = Field([0.0, 90.0], min_length=1)
    wavelengths_nm: list[float] = Field(min_length=1)


class EmissionProblem(_Model):
    """A point dipole near particles in a homogeneous medium: decay-rate (Purcell) and radiative enhancements and
    the efficiency, relative to the same dipole in the medium alone (problems.py:294-309)."""

    kind: Literal["emission"] = "emission"
    particles: list[Particle] = Field(default_factory=list)
    medium_index: float = Field(1.0, gt=0)
    dipole: Dipole = Dipole()
    wavelengths: Wavelengths
    radiation_pattern: RadiationPatternRequest | None = None
    options: dict[str, dict[str, Any]] = Field(default_factory=dict)

    def wavelength_array(self) -> np.ndarray:
        return wavelength_array(self.wavelengths)


def options_for(problem, backend_name: str, method: str) -> dict[str, Any]:
    """Method-level options overlaid with backend-specific ones (problems.py:348-350)."""
    opts = getattr(problem, "options", {}) or {}
    return {**opts.get(method, {}), **opts.get(backend_name, {})}


# --- FDTD options (backends/_fdtd/options.py:28-92) -----------------------------------------------------------------

class FDTDOptions(BaseModel):
    """The reference's FDTDOptions fields. Values this path does not run are accepted by the model and declined by
    the backend: precision "double", grid "auto", mesh_override, symmetry "auto" is run as "off" (full grid)."""

    model_config = ConfigDict(extra="forbid")

    spacing_nm: float | None = Field(None, gt=0)
    cells_per_wavelength: float = Field(20.0, ge=6)
    cells_per_radius: float | None = Field(None, ge=2)
    pml_cells: int | None = Field(None, ge=4)
    pml_kappa: float = Field(1.0, ge=1)
    padding_nm: float | None = Field(None, ge=0)
    pml_gap_nm: float | None = Field(None, ge=0)
    courant: float = Field(0.95, gt=0, le=1)
    run_time_fs: float | None = Field(None, gt=0)
    decay_by: float = Field(1e-5, gt=0, lt=1)
    max_time_fs: float = Field(2000.0, gt=0)
    precision: Literal["auto", "single", "double"] = "auto"
    subpixel: bool = True
    mesh_refinement: Literal["staircase", "volume_average", "subpixel", "conformal"] | None = None
    subpixel_samples: int = Field(8, ge=2, le=24)
    dft_stride: int | None = Field(None, ge=1)
    symmetry: Literal["auto", "off"] = "auto"
    polarization: Literal["te", "tm"] | None = None
    grid: Literal["uniform", "auto"] = "uniform"
    mesh_override: list[dict] = Field(default_factory=list)
    background_cells_per_wavelength: float = Field(60.0, ge=10)
    grading_ratio: float = Field(1.25, ge=1.05, le=1.5)
    cuda_graphs: bool = True
    max_cells: int = Field(300_000_000, ge=1)


def mesh_mode(opts: FDTDOptions) -> str:
    """The interface treatment: mesh_refinement, else conformal (subpixel=True) or staircase (setup.py:101-105)."""
    if opts.mesh_refinement is not None:
        return opts.mesh_refinement
    return "conformal" if opts.subpixel else "staircase"


# --- result (results.py:39-88) ---------------------------------------------------------------------------------------

class NearFieldData(BaseModel):
    """The total field on a plane, arrays indexed [axis1][axis2] (and [component] first for E) (results.py:11-21)."""

    wavelength_nm: float
    plane: str
    offset_nm: float
    axis1_nm: list[float]
    axis2_nm: list[float]
    enhancement: list[list[float]]  # |E|^2 / |E_inc|^2, NaN where undefined
    e_real: list[list[list[float]]] | None = None
    e_imag: list[list[list[float]]] | None = None


class ResonanceData(BaseModel):
    """One resonance, complex frequency omega - i gamma / 2, Q = omega / gamma (results.py:31-37)."""

    wavelength_nm: float  # vacuum wavelength 2 pi c / omega
    energy_ev: float
    q: float
    amplitude: float  # at the probes, relative to the strongest resonance


class FarFieldData(BaseModel):
    wavelength_nm: float
    theta_deg: list[float]
    phi_deg: list[float]
    dcsca_domega_nm2: list[list[float]]  # dC_sca/dOmega in nm^2/sr, [theta][phi]


class Result(BaseModel):
    """Cross sections of a scattering problem."""

    model_config = ConfigDict(extra="forbid")

    kind: Literal["cross_sections"] = "cross_sections"
    method: str
    backend: str
    wavelengths_nm: list[float]
    c_ext_nm2: list[float]
    c_sca_nm2: list[float]
    c_abs_nm2: list[float]
    equivalent_radius_nm: float
    near_field: list[NearFieldData] | None = None
    far_field: list[FarFieldData] | None = None
    resonances: list[ResonanceData] | None = None
    provenance: dict[str, Any] = Field(default_factory=dict)
    warnings: list[str] = Field(default_factory=list)

    @classmethod
    def from_arrays(cls, *, method, backend, wavelengths_nm, c_ext, c_sca, c_abs, r_eq, **kw) -> "Result":
        f = lambda a: np.asarray(a, dtype=float).tolist()  # noqa: E731
        return cls(method=method, backend=backend, wavelengths_nm=f(wavelengths_nm), c_ext_nm2=f(c_ext),
                   c_sca_nm2=f(c_sca), c_abs_nm2=f(c_abs), equivalent_radius_nm=float(r_eq), **kw)

    def _q(self, c) -> np.ndarray:
        return np.asarray(c) / (np.pi * self.equivalent_radius_nm ** 2)

    @property
    def q_ext(self) -> np.ndarray:
        return self._q(self.c_ext_nm2)

    @property
    def q_sca(self) -> np.ndarray:
        return self._q(self.c_sca_nm2)

    @property
    def q_abs(self) -> np.ndarray:
        return self._q(self.c_abs_nm2)


class DiffractionOrder(BaseModel):
    order: tuple[int, int]
    r: dict[str, list[list[float]]]  # reflected power fraction per polarisation, [angle][wavelength]
    t: dict[str, list[list[float]]]


class SpectrumResult(BaseModel):
    """Power reflection, transmission and absorption of layer stacks and periodic structures (results.py:126-175)."""

    model_config = ConfigDict(extra="forbid")

    kind: Literal["spectrum"] = "spectrum"
    method: str
    backend: str
    wavelengths_nm: list[float]
    angles_deg: list[float]
    polarizations: list[str]
    r: dict[str, list[list[float]]]  # total reflectance per polarisation, [angle][wavelength]
    t: dict[str, list[list[float]]]
    a: dict[str, list[list[float]]]
    orders: list[DiffractionOrder] | None = None
    r_amplitude: dict[str, list[list[tuple[float, float]]]] | None = None
    t_amplitude: dict[str, list[list[tuple[float, float]]]] | None = None
    provenance: dict[str, Any] = Field(default_factory=dict)
    warnings: list[str] = Field(default_factory=list)

    def difference(self, other) -> dict[str, dict[str, float]]:
        """Max and RMS absolute difference of r, t, a per polarisation (results.py:176-187; same grids)."""
        if not (np.allclose(self.wavelengths_nm, other.wavelengths_nm)
                and np.allclose(self.angles_deg, other.angles_deg)):
            raise ValueError("spectra are on different wavelength/angle grids")
        out: dict[str, dict[str, float]] = {}
        for q in ("r", "t", "a"):
            mine, theirs = getattr(self, q), getattr(other, q)
            for pol in sorted(set(mine) & set(theirs)):
                d = np.abs(np.asarray(mine[pol]) - np.asarray(theirs[pol]))
                out[f"{q}_{pol}"] = {"max": float(d.max()), "rms": float(np.sqrt(np.mean(d ** 2)))}
        return out


class RadiationPatternData(BaseModel):
    wavelength_nm: float
    theta_deg: list[float]
    phi_deg: list[float]
    dp_domega: list[list[float]]  # radiated power per steradian over P0, [theta][phi]


class EmissionResult(BaseModel):
    """Dipole emission near particles relative to the same dipole in the medium alone (results.py:144-152)."""

    model_config = ConfigDict(extra="forbid")

    kind: Literal["emission"] = "emission"
    method: str
    backend: str
    wavelengths_nm: list[float]
    purcell: list[float]  # total decay-rate enhancement: emitted power / P0
    radiative: list[float]  # power reaching the far field / P0
    efficiency: list[float]  # radiative / purcell
    radiation_pattern: list[RadiationPatternData] | None = None
    provenance: dict[str, Any] = Field(default_factory=dict)
    warnings: list[str] = Field(default_factory=list)


_KINDS = {"scattering": Simulation, "layer_stack": LayerStackProblem, "periodic": PeriodicProblem,
          "emission": EmissionProblem}


def parse_problem(obj):
    """The problem (Simulation, LayerStackProblem or PeriodicProblem, by its ``kind``) from a model (ours, or the
    reference's, through its JSON), a dict or a JSON string."""
    if isinstance(obj, tuple(_KINDS.values())):
        return obj
    if isinstance(obj, BaseModel):
        obj = obj.model_dump(mode="json")
    elif isinstance(obj, str):
        import json
        obj = json.loads(obj)
    kind = obj.get("kind", "scattering") if isinstance(obj, dict) else "scattering"
    if kind not in _KINDS:
        raise ValueError(f"problem kind {kind!r} is not on this path (scattering, layer_stack, periodic, emission)")
    return _KINDS[kind].model_validate(obj)


__all__ = ["Circle", "ConstantIndex", "ConstantPermittivity", "Cylinder", "DiffractionOrder", "Dipole", "Drude",
           "DrudeLorentz", "Ellipse", "EmissionProblem", "EmissionResult", "FarFieldData", "FarFieldRequest", "NearFieldData", "NearFieldRequest", "ResonanceData", "ResonanceRequest", "RadiationPatternData",
           "RadiationPatternRequest", "FDTDOptions", "Incidence", "Layer", "LayerStackProblem", "LayeredSphere", "LorentzPole",
           "Outputs", "Particle", "PatternedLayer", "PeriodicProblem", "PlaneWave", "Polygon", "Rectangle", "Result",
           "Shell", "Simulation", "SpectrumResult", "Sphere", "Spheroid", "Structure2D", "Substrate", "WavelengthRange",
           "mesh_mode", "options_for", "parse_problem", "wavelength_array"]
