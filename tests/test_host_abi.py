"""CPU tests of the C-ABI library surface and of the host-side layout logic (no GPU calls)."""

import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

from lightsmith_amd import _lib
from lightsmith_amd.fdtd import engine as peng

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "lsf.h")


def header_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(lsf_\w+)\s*\(", text)))


def test_library_exports_every_declared_function():
    names = header_functions()
    assert len(names) >= 20
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True, text=True, check=True)
    exported = set(re.findall(r"\bT (lsf_\w+)", out.stdout))
    missing = [n for n in names if n not in exported]
    assert not missing, missing
    assert sorted(_lib.exported_symbols()) == names  # the ctypes table binds exactly the header's functions
    lib = _lib.load()
    for n in names:
        getattr(lib, n)


def test_abi_argument_errors_without_a_gpu():
    lib = _lib.load()
    assert lib.lsf_version().startswith(b"lsf")
    out = ctypes.c_void_p()
    assert lib.lsf_create(None, ctypes.byref(out)) == _lib.LSF_EINVAL
    assert b"null" in lib.lsf_last_error()
    g = _lib.LsfGrid()
    g.shape[:] = [2, 8, 8]
    assert lib.lsf_create(ctypes.byref(g), ctypes.byref(out)) == _lib.LSF_EINVAL
    assert lib.lsf_set_tiling(None, 2, 4, 64, 16) == _lib.LSF_EINVAL
    assert lib.lsf_batch_create(None, 0, ctypes.byref(out)) == _lib.LSF_EINVAL
    assert lib.lsf_batch_step(None, 1, None) == _lib.LSF_EINVAL
    with pytest.raises(ValueError):
        _lib.check(_lib.LSF_EINVAL, "x")
    with pytest.raises(RuntimeError):
        _lib.check(_lib.LSF_EHIP, "x")


def test_grid_struct_matches_the_c_layout():
    # int32 shape[3], complex, periodic[3] | pad | double spacing[3], dt, phase[3][2] | 3 pointers | int32 tick
    assert ctypes.sizeof(_lib.LsfGrid) == 144
    assert _lib.LsfGrid.spacing.offset == 32 and _lib.LsfGrid.dev_E.offset == 112


def test_missing_library_fails_loudly(monkeypatch):
    monkeypatch.setattr(_lib, "_lib", None)
    monkeypatch.setattr(_lib, "LIB_PATH", "/nonexistent/_lsf.so")
    with pytest.raises(ImportError):
        _lib.load()


def test_native_index_round_trip_and_layout():
    shape = (7, 5, 11)
    rng = np.random.default_rng(0)
    ref = rng.integers(0, 3 * np.prod(shape), 500)
    nat = peng.to_native(ref, shape)
    assert np.array_equal(peng.to_reference(nat, shape), ref)
    A = rng.standard_normal((3,) + shape)  # the reference's E[3, N0, N1, N2]
    native = A.transpose(0, 3, 2, 1).reshape(-1)  # component-major, z, y, x fastest
    assert np.array_equal(native[nat], A.reshape(-1)[ref])


@pytest.mark.parametrize("xw", [64, 32])
@pytest.mark.parametrize("shape", [(70, 9, 6), (64, 3, 4), (130, 5, 5), (33, 4, 3)])
def test_ade_chunk_base_reproduces_the_compact_slots(shape, xw):
    """Emulate the E kernel: a wavefront row covers xw x cells; a node's slot = the base of its (plane, row, 32-cell
    chunk, component) + the number of D-form nodes of that component at lower x in the same chunk (the ballot
    restricted to the chunk). It must equal the node's position in slot order, which puts a wavefront row's three
    components' nodes in one contiguous block."""
    rng = np.random.default_rng(1)
    N = int(np.prod(shape))
    nodes = np.sort(rng.choice(3 * N, size=3 * N // 7, replace=False))
    order = np.argsort(peng.ade_slot_key(nodes, shape), kind="stable")
    base = peng.ade_chunk_base(nodes[order], shape)
    n0, n1, n2 = shape
    nchunk = (n0 + 31) // 32
    assert len(base) == 3 * n1 * n2 * nchunk
    c, cell = np.divmod(nodes, N)
    i = cell % n0
    row = cell // n0  # k * n1 + j
    chunk = (row * nchunk + i // 32) * 3 + c
    rank = np.zeros(len(nodes), int)
    for q in np.unique(chunk):
        sel = np.flatnonzero(chunk == q)
        rank[sel] = np.argsort(np.argsort(i[sel]))
    slot = np.empty(len(nodes), int)
    slot[order] = np.arange(len(nodes))
    assert np.array_equal(base[chunk] + rank, slot)
    # a wavefront row's nodes (all components) occupy one contiguous range of slots
    group = row * ((n0 + xw - 1) // xw) + i // xw
    for q in np.unique(group):
        s = np.sort(slot[group == q])
        assert s[-1] - s[0] + 1 == len(s)


def test_material_encoding_is_exact():
    rng = np.random.default_rng(2)
    vals = np.float32(rng.uniform(0.01, 1, 40))
    cb = rng.choice(vals, size=(3, 1000)).astype(np.float32)
    cb[1, ::7] = -2.0  # D-form medium 1
    (ids, tab), n = peng.encode_material(cb)
    assert n <= 41
    for c in range(3):
        assert np.array_equal(tab[(ids >> (8 * c)) & 255], cb[c])
    wide = rng.uniform(0, 1, (3, 400)).astype(np.float32)
    assert peng.encode_material(wide)[0] is None


@pytest.mark.parametrize("cells", [((6, 6), (6, 6), (6, 6)), ((4, 7), (0, 0), (9, 3))])
def test_pml_profiles_match_the_oracle_terms(cells):
    """The per-axis profile tables cover exactly the oracle's CPML slabs (engine.py:388-395), with its b,
    c and kappa at each slab position, for E (node) and H (half-integer) positions."""
    from oracle import fdtd as ofd

    shape, dx, dt = (31, 22, 27), 2.0, 0.9 * 2 / np.sqrt(3)
    spec_o = ofd.PMLSpec(cells=cells, m=3.0, kappa_max=2.5, alpha_max=0.03, eps_ref=1.7)
    spec_p = peng.PMLSpec(cells=cells, m=3.0, kappa_max=2.5, alpha_max=0.03, eps_ref=1.7)
    o = ofd.Engine(shape=shape, dx=dx, dt=dt, inv_eps=np.ones((3,) + shape), pml=spec_o)
    for fld, terms in ((0, o.pml_E), (1, o.pml_H)):
        for d in range(3):
            prof, ns, has_k = peng.pml_axis_profile(spec_p, shape, dt, (dx,) * 3, d, fld)
            mine = {int(p) for p in np.flatnonzero(prof[:, 3] >= 0)}
            theirs = set()
            for c, src, dd, Rs, b, cp, kp, psi in terms:
                if dd != d:
                    continue
                pos = np.arange(Rs[d].start, Rs[d].stop)
                theirs |= set(pos.tolist())
                bb = b.reshape(-1)
                assert np.allclose(prof[pos, 0], bb, rtol=0, atol=1e-15)
                assert (kp is not None) == has_k
            assert mine == theirs
            assert ns == len(mine)
