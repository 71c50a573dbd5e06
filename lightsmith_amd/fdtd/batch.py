"""Many independent FDTD runs stepped together (the oblique sweep's angles x wavelengths x polarisations,
multi-instance design). The reference runs them one Engine after another (periodic.py:503-519); a Batch steps
fully built engines of any shapes with shared launches (lsf_batch_*): one bulk-kernel launch per half-step
and kernel configuration over all members' tiles, fused per-step source operations, one DFT launch for all
monitors, and the whole chunk captured as one HIP graph. Each member evolves exactly as it would alone.
"""

from __future__ import annotations

import ctypes
import itertools
import time

import torch

from .. import _lib
from .._lib import check
from .engine import Engine

# tile shapes with compiled kernels: (rows per lane, wavefronts per block, x cells per wavefront row)
SHAPES = ((1, 4, 64), (2, 4, 64), (1, 8, 64), (2, 8, 64), (1, 4, 32), (2, 4, 32), (1, 8, 32), (2, 8, 32))


class Batch:
    """Members share one DFT stride; they keep their own step counters, sources and monitors."""

    def __init__(self, engines, schedule=None):
        self.engines = list(engines)
        if not self.engines:
            raise ValueError("a batch needs at least one engine")
        for e in self.engines:
            if not isinstance(e, Engine):
                raise ValueError("batch members are lightsmith_amd Engines")
        strides = {e.stride for e in self.engines}
        if len(strides) != 1:
            raise ValueError(f"batch members must share one DFT stride, got {sorted(strides)}")
        self.stride = strides.pop()
        devs = {str(e._dev) for e in self.engines}
        if len(devs) != 1:
            raise ValueError("batch members must live on one device")
        self._dev = self.engines[0]._dev
        self.lib = _lib.load()
        arr = (ctypes.c_void_p * len(self.engines))(*[e._h.value for e in self.engines])
        h = ctypes.c_void_p()
        check(self.lib.lsf_batch_create(arr, len(self.engines), ctypes.byref(h)), "lsf_batch_create")
        self._h = h
        for e in self.engines:
            e._held += 1
        self._graphed = False
        self.schedule = None
        self.set_schedule(*(schedule or (2, 4, 64, 16, 0)))

    # --- schedule ------------------------------------------------------------------------------------
    def set_schedule(self, ry: int, waves: int, xw: int, zc: int, grid_blocks: int = 0):
        """Tile shape of all members' bulk kernels and grid: grid_blocks = 0 launches one block per tile,
        grid_blocks > 0 a persistent grid of that many blocks striding over the tiles."""
        check(self.lib.lsf_batch_set_schedule(self._h, int(ry), int(waves), int(xw), int(zc), int(grid_blocks)),
              "lsf_batch_set_schedule")
        self.schedule = (int(ry), int(waves), int(xw), int(zc), int(grid_blocks))

    def tiles(self) -> int:
        n = ctypes.c_int64()
        check(self.lib.lsf_batch_tiles(self._h, ctypes.byref(n)), "lsf_batch_tiles")
        return int(n.value)

    @property
    def cells(self) -> int:
        return sum(e.N for e in self.engines)

    def state_tensors(self) -> list:
        out = []
        for e in self.engines:
            out += e.state_tensors()
        return out

    def candidates(self, persistent: bool = True):
        """Schedules the autotuner tries: every compiled tile shape x z chunks x (one block per tile, or a
        persistent grid of 2 and 4 blocks per CU)."""
        cus = torch.cuda.get_device_properties(self._dev).multi_processor_count
        grids = [0] + ([2 * cus, 4 * cus] if persistent else [])
        return [(ry, w, xw, zc, gb) for (ry, w, xw), zc, gb in itertools.product(SHAPES, (8, 16, 32), grids)]

    def autotune(self, candidates=None, steps: int = 10, warmup: int = 2, keep: bool = True) -> list:
        """Time the bulk kernels (HIP events) of each candidate schedule over `steps` steps, keep the fastest
        and restore every member's state afterwards (the trial steps leave no trace). Returns
        [(ms per step, schedule)] sorted, fastest first."""
        candidates = list(candidates or self.candidates())
        snap = [t.clone() for t in self.state_tensors()]
        stream = self._stream()
        results = []
        try:
            for cand in candidates:
                self.set_schedule(*cand)
                check(self.lib.lsf_batch_step(self._h, warmup, stream), "lsf_batch_step")
                ms = (ctypes.c_double * 2)(0.0, 0.0)
                check(self.lib.lsf_batch_step_timed(self._h, steps, stream, ms), "lsf_batch_step_timed")
                results.append(((ms[0] + ms[1]) / steps, tuple(cand)))
        finally:
            torch.cuda.synchronize(self._dev)
            for t, s in zip(self.state_tensors(), snap):
                t.copy_(s)
            torch.cuda.synchronize(self._dev)
        results.sort()
        self.set_schedule(*(results[0][1] if keep else self.schedule))
        return results

    # --- stepping --------------------------------------------------------------------------------------
    def _stream(self):
        return ctypes.c_void_p(torch.cuda.current_stream(self._dev).cuda_stream)

    def _advance(self, n: int):
        for e in self.engines:
            e.n_host += n

    def step(self, n: int = 1):
        check(self.lib.lsf_batch_step(self._h, int(n), self._stream()), "lsf_batch_step")
        self._advance(int(n))

    def steps(self, n: int):
        self.step(n)

    def step_timed(self, n: int) -> tuple[float, float]:
        """n steps; returns the summed (H, E) bulk-kernel times in ms (synchronises)."""
        ms = (ctypes.c_double * 2)(0.0, 0.0)
        check(self.lib.lsf_batch_step_timed(self._h, int(n), self._stream(), ms), "lsf_batch_step_timed")
        self._advance(int(n))
        return ms[0], ms[1]

    def sample(self):
        check(self.lib.lsf_batch_sample(self._h, self._stream()), "lsf_batch_sample")

    def chunk(self):
        check(self.lib.lsf_batch_chunk(self._h, 1, self.stride, self._stream()), "lsf_batch_chunk")
        self._advance(self.stride)

    def make_runner(self, use_graphs: bool = True):
        """(run_one_chunk, used_graphs), as Engine.make_runner: the first chunk runs eagerly, then one chunk of
        all members is captured as a single HIP graph and replayed."""
        if not use_graphs:
            return self.chunk, False
        stream = torch.cuda.Stream(device=self._dev)
        stream.wait_stream(torch.cuda.current_stream(self._dev))
        state = {"captured": False}
        self._graph_stream = stream

        def run():
            if not state["captured"]:
                self.chunk()
                torch.cuda.current_stream(self._dev).synchronize()
                check(self.lib.lsf_batch_graph_capture(self._h, self.stride, ctypes.c_void_p(stream.cuda_stream)),
                      "lsf_batch_graph_capture")
                state["captured"] = True
                self._graphed = True
                return
            stream.wait_stream(torch.cuda.current_stream(self._dev))
            check(self.lib.lsf_batch_graph_replay(self._h, 1, ctypes.c_void_p(stream.cuda_stream)),
                  "lsf_batch_graph_replay")
            torch.cuda.current_stream(self._dev).wait_stream(stream)
            self._advance(self.stride)

        return run, True

    def release(self):
        if self._graphed:
            check(self.lib.lsf_batch_graph_reset(self._h), "lsf_batch_graph_reset")
            self._graphed = False

    def run(self, n_chunks: int, use_graphs: bool = True) -> float:
        """n_chunks chunks of every member; returns the wall time (s) between two device synchronisations."""
        run, _ = self.make_runner(use_graphs)
        torch.cuda.synchronize(self._dev)
        t0 = time.perf_counter()
        for _ in range(n_chunks):
            run()
        torch.cuda.synchronize(self._dev)
        return time.perf_counter() - t0

    def close(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            torch.cuda.synchronize(self._dev)
            self.lib.lsf_batch_destroy(h)
            self._h = ctypes.c_void_p()
            for e in self.engines:
                e._held -= 1

    def __del__(self):
        try:
            self.close()
        except Exception:  # noqa: BLE001 (interpreter shutdown)
            pass
