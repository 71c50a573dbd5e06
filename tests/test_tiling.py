"""The solve-level drivers' launch tiling (engine.filling_tiling): the longest z chunk of 8, 4 or 2 that still gives
16 row marches per SIMD, 32-cell rows for cells no wider than 32. The picks below are the fastest chunks measured on
the throughput scene (96^3 and 128^3: 2; 160^3 and 192^3: 4; 256^3 and 320^3: 8) and on the 256 x 192 x 96 adjoint
scene (4: forward 299 -> 288 ms against 8)."""

import pytest

from lightsmith_amd.fdtd.engine import filling_tiling


@pytest.mark.parametrize("n, zc", [(96, 2), (97, 2), (128, 2), (160, 4), (192, 4), (256, 8), (320, 8)])
def test_cubic_grids_take_the_measured_z_chunk(n, zc):
    assert filling_tiling([(n, n, n)]) == (1, 4, 64, zc)


def test_adjoint_scene_and_narrow_cells():
    assert filling_tiling([(256, 192, 96)]) == (1, 4, 64, 4)
    assert filling_tiling([(20, 20, 300)] * 4) == (1, 4, 32, 2)  # periodic unit cells: 32-cell rows
    assert filling_tiling([(33, 20, 300)]) == (1, 4, 64, 2)


def test_members_fill_the_gpu_together():
    small = (97, 97, 97)
    assert filling_tiling([small]) == (1, 4, 64, 2)
    assert filling_tiling([small] * 16) == (1, 4, 64, 8)  # 16 such grids stepped as one batch
