"""Build the HIP library in-tree: ``python -m lightsmith_amd.build``.

Translation units (compiled by hipcc for gfx950 in parallel, then linked into lightsmith_amd/_lsf.so):
  csrc/lsf.hip     host API, batches, sparse per-step kernels
  csrc/lsf_k_adj.hip the adjoint bulk kernels
  csrc/lsf_k_fused.hip the single-pass step kernel
  csrc/lsf_k_h.hip the bulk H kernel instantiations, once per field type (float, complex64)
  csrc/lsf_k_e.hip the bulk E kernel instantiations, once per field type, material mode (dense, indexed)
                   and periodicity
Packed-FP32 instructions are switched off (-packed-fp32-ops): the kernels are bandwidth-bound and the
scalar forms are the ones this project trusts.
Floating-point contraction is off (-ffp-contract=off): the kernels fuse multiply-adds explicitly (fmad in
csrc/lsf_common.h), so every instantiation of a kernel rounds alike.
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys
import tempfile
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "_lsf.so")
ARCH = "gfx950"

# (object name, source, extra defines)
UNITS = [("lsf", "lsf.hip", []), ("k_adj", "lsf_k_adj.hip", []), ("k_fused", "lsf_k_fused.hip", [])] + [
    (f"k_h_{t}", "lsf_k_h.hip", [f"-DLSF_H_CPLX={1 if t == 'c' else 0}", f"-DLSF_H_ENTRY=launch_update_h_{t}"])
    for t in ("f", "c")
] + [
    (f"k_e_{t}{m}{p}", "lsf_k_e.hip", [f"-DLSF_E_CPLX={1 if t == 'c' else 0}", f"-DLSF_E_MAT={m}", f"-DLSF_E_PER={p}",
                                       f"-DLSF_E_ENTRY=launch_update_e_{t}{m}{p}"])
    for t in ("f", "c") for m in (0, 1) for p in (0, 1)
]


def hipcc() -> str:
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found (ROCm is required to build lightsmith_amd)")


def flags() -> list[str]:
    return [f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-ffp-contract=off",
            "-Wno-unused-command-line-argument",
            "-I", os.path.join(ROOT, "include"), "-I", CSRC]


def compile_cmd(src: str, obj: str, defines: list[str]) -> list[str]:
    return [hipcc(), *flags(), *defines, "-c", os.path.join(CSRC, src), "-o", obj]


def link_cmd(objs: list[str], out: str) -> list[str]:
    return [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", out]


def _run(cmd: list[str]) -> subprocess.CompletedProcess:
    return subprocess.run(cmd, capture_output=True, text=True)


def build(verbose: bool = False, jobs: int | None = None) -> str:
    """Compile the library (always; the build is the check that the sources compile)."""
    jobs = jobs or min(len(UNITS), max(1, min(16, os.cpu_count() or 1)))
    with tempfile.TemporaryDirectory(prefix="lsf_build_") as tmp:
        cmds = [compile_cmd(src, os.path.join(tmp, name + ".o"), defs) for name, src, defs in UNITS]
        with ThreadPoolExecutor(max_workers=jobs) as pool:
            procs = list(pool.map(_run, cmds))
        for cmd, proc in zip(cmds, procs):
            if proc.returncode != 0:
                sys.stderr.write(proc.stdout + proc.stderr)
                raise RuntimeError(f"hipcc failed ({proc.returncode}): {' '.join(cmd)}")
            if verbose:
                sys.stderr.write(proc.stderr)
        out_tmp = OUT + ".tmp"
        cmd = link_cmd([c[-1] for c in cmds], out_tmp)
        proc = _run(cmd)
        if proc.returncode != 0:
            sys.stderr.write(proc.stdout + proc.stderr)
            raise RuntimeError(f"link failed ({proc.returncode}): {' '.join(cmd)}")
    os.replace(out_tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
