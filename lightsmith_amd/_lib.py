"""ctypes binding of the HIP library ``_lsf.so`` (declarations: include/lsf.h).

The library must be built in-tree (``__graft_entry__.build()`` or ``python -m lightsmith_amd.build``).
There is no fallback: a missing library raises ImportError on first use. torch is imported before the
library is loaded so that both share torch's HIP runtime (one ``libamdhip64.so.7`` per process).
"""

from __future__ import annotations

import ctypes
import os

import torch  # noqa: F401  (loads the HIP runtime the library binds to)

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_lsf.so")

c_int, c_i64, c_vp, c_dbl, c_char_p = ctypes.c_int, ctypes.c_int64, ctypes.c_void_p, ctypes.c_double, ctypes.c_char_p

LSF_OK, LSF_EINVAL, LSF_EHIP, LSF_ESTATE = 0, 1, 2, 3
PH_CORRECT_H, PH_ADVANCE_H, PH_OFFDIAG, PH_CORRECT_E, PH_GROUPS, PH_ADVANCE_E = range(6)


class LsfGrid(ctypes.Structure):
    _fields_ = [("shape", ctypes.c_int32 * 3), ("complex_fields", ctypes.c_int32), ("periodic", ctypes.c_int32 * 3),
                ("spacing", c_dbl * 3), ("dt", c_dbl), ("bloch_phase", (c_dbl * 2) * 3), ("dev_E", c_vp),
                ("dev_H", c_vp), ("dev_counter", c_vp), ("tick", ctypes.c_int32)]


# name -> argument types (every function returns int unless listed in _RESTYPES)
_SIGNATURES = {
    "lsf_create": [ctypes.POINTER(LsfGrid), ctypes.POINTER(c_vp)],
    "lsf_destroy": [c_vp],
    "lsf_set_material_dense": [c_vp, c_vp],
    "lsf_set_material_indexed": [c_vp, c_vp, c_vp],
    "lsf_set_pml": [c_vp, c_int, c_int, c_vp, c_int, c_vp, c_vp, c_vp, c_int],
    "lsf_set_ade": [c_vp, c_vp, c_vp, c_i64, c_int, c_vp, c_int],
    "lsf_add_gather": [c_vp, c_int, c_int, c_i64, c_int, c_vp, c_vp, c_vp, c_vp, c_vp],
    "lsf_add_modal": [c_vp, c_int, c_int, c_i64, c_int, c_vp, c_vp, c_vp, c_i64, c_vp],
    "lsf_add_table": [c_vp, c_int, c_int, c_i64, c_vp, c_vp, c_vp, c_int, c_i64, c_vp, c_vp, c_vp, c_vp, c_vp],
    "lsf_add_table_view": [c_vp, c_int, c_int, c_vp, c_vp, c_int, c_i64, c_vp, c_vp, c_vp],
    "lsf_add_line1d": [c_vp, c_int, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_int, c_vp, c_i64],
    "lsf_add_subengine": [c_vp, c_vp],
    "lsf_add_mixed_group": [c_vp, c_i64, c_int, c_int, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_int, c_int],
    "lsf_add_dft_monitor": [c_vp, c_int, c_i64, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_int, c_int],
    "lsf_add_time_monitor": [c_vp, c_i64, c_vp, c_vp, c_i64, c_vp],
    "lsf_set_energy": [c_vp, c_vp],
    "lsf_set_tiling": [c_vp, c_int, c_int, c_int, c_int],
    "lsf_fuse_table_tensor": [c_vp],
    "lsf_set_tile_order": [c_vp, c_int],
    "lsf_step": [c_vp, c_int, c_vp],
    "lsf_step_noalloc": [c_vp, c_int, c_vp],
    "lsf_step_timed": [c_vp, c_int, c_vp, ctypes.POINTER(c_dbl)],
    "lsf_sample": [c_vp, c_vp],
    "lsf_chunk": [c_vp, c_int, c_int, c_vp],
    "lsf_graph_capture": [c_vp, c_int, c_vp],
    "lsf_graph_replay": [c_vp, c_int, c_vp],
    "lsf_graph_reset": [c_vp],
    "lsf_add_tensor_group": [c_vp, c_i64, c_vp, c_vp, c_i64, c_vp, c_vp, c_vp, c_vp, c_int, c_int, c_vp, c_vp, c_vp,
                             c_vp, c_vp],
    "lsf_set_fused": [c_vp, c_vp, c_vp, ctypes.POINTER(c_vp)],
    "lsf_parity": [c_vp],
    "lsf_set_adjoint": [c_vp, c_int],
    "lsf_add_tape": [c_vp, c_i64, c_vp, c_vp, c_i64, c_vp],
    "lsf_add_adjoint_gradient": [c_vp, c_i64, c_vp, c_vp, c_vp, c_i64, c_vp, c_i64, c_vp],
    "lsf_add_adjoint_ade": [c_vp, c_i64, c_vp, c_vp, c_int, c_int, c_vp, c_vp, c_vp],
    "lsf_add_adjoint_source": [c_vp, c_int, c_int, c_i64, c_vp, c_dbl, c_vp, c_i64, c_vp, c_vp, c_vp, c_vp, c_i64,
                               c_int],
    "lsf_batch_create": [ctypes.POINTER(c_vp), c_int, ctypes.POINTER(c_vp)],
    "lsf_batch_destroy": [c_vp],
    "lsf_batch_set_schedule": [c_vp, c_int, c_int, c_int, c_int, c_int],
    "lsf_batch_tiles": [c_vp, ctypes.POINTER(c_i64)],
    "lsf_batch_step": [c_vp, c_int, c_vp],
    "lsf_batch_step_timed": [c_vp, c_int, c_vp, ctypes.POINTER(c_dbl)],
    "lsf_batch_sample": [c_vp, c_vp],
    "lsf_batch_chunk": [c_vp, c_int, c_int, c_vp],
    "lsf_batch_graph_capture": [c_vp, c_int, c_vp],
    "lsf_batch_graph_replay": [c_vp, c_int, c_vp],
    "lsf_batch_graph_reset": [c_vp],
    "lsf_last_error": [],
    "lsf_version": [],
}
_RESTYPES = {"lsf_destroy": None, "lsf_batch_destroy": None, "lsf_last_error": c_char_p, "lsf_version": c_char_p}

_lib = None


def load() -> ctypes.CDLL:
    """The loaded library (ImportError when it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"lightsmith_amd: HIP library {LIB_PATH} is not built; run "
                              "`python -c 'import __graft_entry__ as g; g.build()'` or `python -m lightsmith_amd.build`")
        lib = ctypes.CDLL(LIB_PATH)
        for name, args in _SIGNATURES.items():
            fn = getattr(lib, name)
            fn.argtypes = args
            fn.restype = _RESTYPES.get(name, c_int)
        _lib = lib
    return _lib


def exported_symbols() -> list[str]:
    return list(_SIGNATURES)


def check(rc: int, what: str = "") -> None:
    """Map a library status to the reference's exception types (setup.py:753 uses FloatingPointError for
    divergence; configuration errors are ValueError, runtime errors RuntimeError)."""
    if rc == LSF_OK:
        return
    msg = load().lsf_last_error().decode(errors="replace")
    text = f"{what}: {msg}" if what else msg
    if rc == LSF_EINVAL:
        raise ValueError(text)
    raise RuntimeError(text)


def ptr(t) -> int | None:
    """Device pointer of a tensor (None for None)."""
    return None if t is None else t.data_ptr()
