"""Many independent evolutions of uneven grid sizes in shared launches per
RK4 stage (one workgroup per tile of any instance), with HIP-graph-captured
steps and an autotuner that picks the tile shape and the launch size.

The reference evolves one grid per call (evolution/time_evolve.py:144-217);
``BatchEvolver.step`` applies the same ``rk4_step`` to every instance of the
batch.  Instances may have periodic or Sommerfeld faces (Sommerfeld faces on
the tile grid of the tile shape used: extents that are multiples of 4 or 8
along those axes), mad_q = 1 and kappa = 0; other configurations run
through ``Evolver`` one grid at a time.
"""

import ctypes
import time

import torch

from . import _lib
from .time_evolve import precision_code, stream_handle
from .variables import BSSNVariables, device, pack, unpack

TILE_VARIANTS = {0: (4, 8, 8), 1: (8, 8, 4), 2: (8, 4, 8)}
_TUNED = {}  # (shapes, precision) -> (tile_variant, grid_blocks)


def _cu_count():
    return torch.cuda.get_device_properties(torch.cuda.current_device()).multi_processor_count


class BatchEvolver:
    """States resident in HBM, one packed (21, nx, ny, nz) tensor each."""

    def __init__(self, states, params, precision="f64", tile_variant=None, grid_blocks=None, autotune=True):
        if len(states) != len(params) or not states:
            raise ValueError("one parameter tuple per state is required")
        dev = device()
        self.states = [(pack(s) if isinstance(s, BSSNVariables) else s).to(dev, torch.float64).contiguous().clone()
                       for s in states]
        self.params = list(params)
        self.precision = precision
        self.n = len(self.states)
        self.shapes = tuple(tuple(s.shape[1:]) for s in self.states)
        grids = (_lib.CGrid * self.n)(*[_lib.c_grid(s) for s in self.shapes])
        self._grids = grids
        self._cparams = (_lib.CParams * self.n)(*[_lib.c_params(p) for p in self.params])
        lib = _lib.load()
        size = int(lib.bssn_batch_work_size(grids, self.n))
        if size < 0:
            _lib.check(size)
        self.work = torch.empty(size, dtype=torch.float64, device=dev)
        self.plan = _lib.CBatchPlan()
        self.graph = None
        key = (self.shapes, tuple((p.xl_bc, p.xr_bc, p.yl_bc, p.yr_bc, p.zl_bc, p.zr_bc) for p in self.params),
               precision)
        self.tuning = {"source": "given"}
        if tile_variant is None or grid_blocks is None:
            if autotune and key in _TUNED:
                tile_variant, grid_blocks = _TUNED[key]
                self.tuning = {"source": "cached"}
            elif autotune:
                tile_variant, grid_blocks = self.autotune()
                _TUNED[key] = (tile_variant, grid_blocks)
            else:
                tile_variant, grid_blocks = 0, 1 << 20
                self.tuning = {"source": "default"}
        self.prepare(tile_variant, grid_blocks)
        self.tuning.update(tile_variant=self.tile_variant, tile_shape=TILE_VARIANTS[self.tile_variant],
                           grid_blocks=self.grid_blocks)

    @property
    def cells(self):
        return sum(int(torch.tensor(s).prod()) for s in self.shapes)

    def prepare(self, tile_variant, grid_blocks):
        lib = _lib.load()
        ptrs = (ctypes.c_void_p * self.n)(*[s.data_ptr() for s in self.states])
        _lib.check(lib.bssn_batch_prepare(self._cparams, self._grids, ptrs, self.n, precision_code(self.precision),
                                          int(tile_variant), int(grid_blocks), ctypes.c_void_p(self.work.data_ptr()),
                                          ctypes.byref(self.plan), stream_handle()))
        self.tile_variant, self.grid_blocks = int(tile_variant), int(grid_blocks)
        self.graph = None

    def _step_once(self):
        _lib.check(_lib.load().bssn_batch_step(ctypes.byref(self.plan), ctypes.c_void_p(self.work.data_ptr()),
                                               stream_handle()))

    def step(self, nsteps=1):
        """``nsteps`` RK4 steps of every instance (graph replays once captured)."""
        for _ in range(int(nsteps)):
            if self.graph is not None:
                self.graph.replay()
            else:
                self._step_once()
        return self

    def capture(self):
        """Capture one batched RK4 step in a HIP graph; ``step`` then replays it.
        The warm-up step that precedes the capture advances the states by one
        step (the captured launches themselves only run on replay)."""
        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s):
            self._step_once()  # warm-up on the capture stream
        torch.cuda.current_stream().wait_stream(s)
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            self._step_once()
        self.graph = g
        return self

    def autotune(self, reps=2):
        """Time every tile shape on scratch copies of the states, with each
        tile list in one launch and in launches of 16 workgroups per CU;
        return the fastest (tile_variant, grid_blocks).  Tile shapes whose
        grid some instance's Sommerfeld faces are not on are skipped."""
        saved = [s.clone() for s in self.states]
        cus = _cu_count()
        best, best_t = None, float("inf")
        for v in TILE_VARIANTS:
            for blocks in (16 * cus, 1 << 20):
                try:
                    self.prepare(v, blocks)
                except NotImplementedError:
                    break
                self._step_once()
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                for _ in range(reps):
                    self._step_once()
                torch.cuda.synchronize()
                t = (time.perf_counter() - t0) / reps
                if t < best_t:
                    best, best_t = (v, blocks), t
        for s, v in zip(self.states, saved):
            s.copy_(v)
        if best is None:
            raise NotImplementedError("no tile shape of the batched launch fits the Sommerfeld faces of this batch")
        self.tuning = {"source": "autotuned", "step_seconds": best_t}
        return best

    def variables(self, like=None):
        return [unpack(s, like) for s in self.states]
