"""FDTD time stepping on MI355X: the engine, its sources and monitors, and run control."""

from .engine import (DFTMonitor, DispersiveGroup, Engine, OffDiagonal, PMLSpec, TensorMesh, TimeMonitor, e_region,
                     flat_index, h_region)
from .media import HBARC_EV_NM, Medium, Pole, discrete_eps, drude_lorentz, pole_coefficients, time_step
from .sources import BlochPlaneSource, ModeInjection, PointDipole, TableSource, TFSFPlaneWave, tfsf_slabs

__all__ = ["Engine", "PMLSpec", "DFTMonitor", "TimeMonitor", "DispersiveGroup", "OffDiagonal", "TensorMesh",
           "flat_index",
           "e_region", "h_region", "Medium", "Pole", "HBARC_EV_NM", "drude_lorentz", "pole_coefficients",
           "discrete_eps", "time_step", "PointDipole", "TFSFPlaneWave", "TableSource", "BlochPlaneSource",
           "ModeInjection", "tfsf_slabs"]
