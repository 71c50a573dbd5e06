"""lightsmith_amd: lightsmith's FDTD time-stepping hot path, MI355X-native.

Python host code on PyTorch-ROCm (device memory and streams) drives hand-written gfx950 HIP kernels
through the C ABI declared in include/lsf.h. Entry points mirror the reference: ``fdtd.Engine`` (the
engine.py contract), its sources and monitors, and ``fdtd.run_until_decayed``.
"""

__version__ = "0.1.0"
