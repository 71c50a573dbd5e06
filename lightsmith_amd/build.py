"""Build the HIP library in-tree: ``python -m lightsmith_amd.build``.

One translation unit (csrc/lsf.hip) compiled by hipcc for gfx950 into lightsmith_amd/_lsf.so.
Packed-FP32 instructions are switched off (-packed-fp32-ops): the kernels are bandwidth-bound and
the scalar forms are the ones this project trusts.
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SRC = os.path.join(HERE, "csrc", "lsf.hip")
OUT = os.path.join(HERE, "_lsf.so")
ARCH = "gfx950"


def hipcc() -> str:
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found (ROCm is required to build lightsmith_amd)")


def command(out: str = OUT) -> list[str]:
    return [hipcc(), f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-Wno-unused-command-line-argument",
            "-I", os.path.join(ROOT, "include"), "-I", os.path.join(HERE, "csrc"), SRC, "-o", out]


def build(verbose: bool = False) -> str:
    """Compile the library (always; the build is the check that the sources compile)."""
    tmp = OUT + ".tmp"
    cmd = command(tmp)
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        sys.stderr.write(proc.stdout + proc.stderr)
        raise RuntimeError(f"hipcc failed ({proc.returncode}): {' '.join(cmd)}")
    if verbose:
        sys.stderr.write(proc.stderr)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
