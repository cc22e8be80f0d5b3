"""Tiled evolution on one GPU: the reference's parallel/ path with the
multi-device halo exchange turned into an intra-GPU ghost fill.

Names follow reference parallel/: ``TileLayout`` (layout.py:27-97),
``require_bssn_guards`` (:81-97), ``local_parameters`` (evolution.py:19-38),
``initialize_tiled_bssn`` (initialization.py:86-114),
``refresh_tiled_field`` / ``refresh_bssn_halos`` (communication.py:142-147,
evolution.py:74-92), ``compute_tiled_bssn_rhs`` (:105-108),
``tiled_rk4_step`` (:123-142) and ``tiled_constraint_norms`` (:173-191).

Reference-shaped tiled fields have the components first, then the tile grid,
then the allocated tile (``comp + counts + allocated_shape``).  Internally a
tiled state is one device buffer, tile-major, each tile a packed 21-component
state of the allocated extent, so every tile runs the ordinary fused stage
kernels with its tile-local parameters (guard cells, owned range, owned
faces, shifted origin) and ``bssn_tiled_ghost_fill`` refreshes every guard
after each stage.
"""

import ctypes
from typing import NamedTuple

import numpy as np
import torch

from . import _lib
from .constraints import NORM_NAMES
from .time_evolve import _ptr, precision_code, require_device_f64, stream_handle
from .variables import NCOMP, BSSNVariables, device

FACE_NAMES = ("xl_bc", "xr_bc", "yl_bc", "yr_bc", "zl_bc", "zr_bc")
FACE_NAME_PAIRS = (("xl_bc", "xr_bc"), ("yl_bc", "yr_bc"), ("zl_bc", "zr_bc"))
EXTERIOR = {"linear": 0, "constant": 1, "preserve": 2}
_COMPONENTS = ((3, 3), (), (3, 3), (), (3,), (), (3,))
_SYM = ((0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2))


class TileLayout(NamedTuple):
    """Tile grid, owned tile extent, guard depth and owned upper endpoints."""

    counts: tuple
    tile_shape: tuple
    guard_cells: int = 4
    upper_endpoints: tuple = (False, False, False)

    @property
    def ntiles(self):
        return int(np.prod(self.counts))

    @property
    def active_shape(self):
        return tuple(w * c + int(e) for w, c, e in zip(self.tile_shape, self.counts, self.upper_endpoints))

    @property
    def allocated_shape(self):
        return tuple(w + 2 * self.guard_cells for w in self.tile_shape)

    def tiles(self):
        c = self.counts
        return [(i, j, k) for i in range(c[0]) for j in range(c[1]) for k in range(c[2])]

    def validate(self):
        if self.guard_cells < 1:
            raise ValueError("Tiled fields require at least one guard cell")
        for w, c, e in zip(self.tile_shape, self.counts, self.upper_endpoints):
            if w < 1:
                raise ValueError("Tiled fields require at least one owned sample per axis")
            if c > 1 and w < self.guard_cells:
                raise ValueError("A tile with neighbors must be at least as wide as its halo")
            if e and self.guard_cells < 2:
                raise ValueError("An owned upper endpoint consumes a guard slot, so its axis needs at least two")

    def c_layout(self):
        return _lib.CTileLayout((ctypes.c_int32 * 3)(*self.counts), (ctypes.c_int32 * 3)(*self.tile_shape),
                                int(self.guard_cells), (ctypes.c_int32 * 3)(*(int(e) for e in self.upper_endpoints)))


def require_bssn_guards(layout, params):
    """Guard depth of the MAD / KO stencils (reference layout.py:81-97)."""
    if layout.guard_cells < 4:
        raise ValueError("Tiled BSSN evolution requires four guards for the supported MAD stencils")
    if layout.guard_cells >= 5 or float(params.mad_q) == 1.0:
        return
    for endpoint, names in zip(layout.upper_endpoints, FACE_NAME_PAIRS):
        if endpoint and all(int(getattr(params, n)) != 1 for n in names):
            raise ValueError("General MAD with an owned periodic upper endpoint requires five guard cells")


def local_parameters(params, layout, idx):
    """Coordinates, owned bounds and physical-face flags of tile ``idx``."""
    g = layout.guard_cells
    origins = (params.x_min, params.y_min, params.z_min)
    minima = tuple(o + (t * w - g) * params.dx for o, t, w in zip(origins, idx, layout.tile_shape))
    owned = tuple(w + int(e and t == c - 1) for w, t, c, e in
                  zip(layout.tile_shape, idx, layout.counts, layout.upper_endpoints))
    faces = tuple(f for t, c in zip(idx, layout.counts) for f in (t == 0, t == c - 1))
    reduced = params.reduced_axes if params.reduced_axes is not None else tuple(n == 1 for n in layout.active_shape)
    return params._replace(x_min=minima[0], y_min=minima[1], z_min=minima[2], guard_cells=g, owned_shape=owned,
                           owned_faces=faces, reduced_axes=reduced)


def _owned_slices(layout, idx):
    g = layout.guard_cells
    owned = tuple(w + int(e and t == c - 1) for w, t, c, e in
                  zip(layout.tile_shape, idx, layout.counts, layout.upper_endpoints))
    glob = tuple(slice(t * w, t * w + n) for t, w, n in zip(idx, layout.tile_shape, owned))
    loc = tuple(slice(g, g + n) for n in owned)
    return glob, loc


def _bcs(params):
    return (ctypes.c_int32 * 6)(*(int(getattr(params, n)) for n in FACE_NAMES))


def refresh_tiled_field(field, params, layout, exterior="linear"):
    """In-place ghost fill of a tile-major device field (ntiles, ncomp, *alloc)."""
    if exterior not in EXTERIOR:
        raise ValueError("Unknown physical exterior rule")
    require_device_f64(field, "tiled field")
    want = (layout.ntiles,) + tuple(layout.allocated_shape)
    if field.dim() != 5 or field.shape[0] != want[0] or tuple(field.shape[2:]) != want[1:]:
        raise ValueError(f"tiled field has shape {tuple(field.shape)}; expected ({want[0]}, ncomp) + {want[1:]}")
    cl = layout.c_layout()
    _lib.check(_lib.load().bssn_tiled_ghost_fill(ctypes.byref(cl), _bcs(params), EXTERIOR[exterior],
                                                 int(field.shape[1]), _ptr(field), stream_handle()))
    return field


# ------------------------------------------------------------ conversions

def tile_global_packed(packed, layout):
    """Guard-free global (ncomp, *active) device tensor -> tile-major
    (ntiles, ncomp, *alloc) with the owned samples set and guards zero."""
    out = torch.zeros((layout.ntiles, packed.shape[0]) + layout.allocated_shape, dtype=torch.float64,
                      device=packed.device)
    for t, idx in enumerate(layout.tiles()):
        glob, loc = _owned_slices(layout, idx)
        out[(t, slice(None)) + loc] = packed[(slice(None),) + glob]
    return out


def gather_owned(tiled, layout):
    """Tile-major (ntiles, ncomp, *alloc) -> global (ncomp, *active)."""
    out = torch.empty((tiled.shape[1],) + layout.active_shape, dtype=tiled.dtype, device=tiled.device)
    for t, idx in enumerate(layout.tiles()):
        glob, loc = _owned_slices(layout, idx)
        out[(slice(None),) + glob] = tiled[(t, slice(None)) + loc]
    return out


def _to_tensor(x, dev):
    if isinstance(x, torch.Tensor):
        return x.to(device=dev, dtype=torch.float64)
    return torch.as_tensor(np.asarray(x, dtype=np.float64), device=dev)


def from_reference_tiled(vars, layout):
    """Reference-shaped tiled BSSNVariables -> tile-major packed device tensor."""
    dev = device()
    comps = []
    for f, lead in zip(vars, _COMPONENTS):
        t = _to_tensor(f, dev)
        if lead == (3, 3):
            comps.extend(t[i, j] for i, j in _SYM)
        elif lead == (3,):
            comps.extend(t[i] for i in range(3))
        else:
            comps.append(t)
    packed = torch.stack(comps)  # (21, cx, cy, cz, *alloc)
    packed = packed.reshape((NCOMP, layout.ntiles) + layout.allocated_shape)
    return packed.transpose(0, 1).contiguous()


def to_reference_tiled(tiled, layout, like=None):
    """Tile-major packed tensor -> reference-shaped tiled BSSNVariables."""
    p = tiled.transpose(0, 1).reshape((NCOMP,) + tuple(layout.counts) + layout.allocated_shape)

    def sym(first):
        rows = [torch.stack([p[first + _SYM.index((min(i, j), max(i, j)))] for j in range(3)]) for i in range(3)]
        return torch.stack(rows)
    fields = BSSNVariables(sym(0), p[6].clone(), sym(7), p[13].clone(), p[14:17].clone(), p[17].clone(),
                           p[18:21].clone())
    if like is None or isinstance(like, torch.Tensor):
        return fields
    return BSSNVariables(*(f.cpu().numpy() for f in fields))


def _pack_global(vars, dev):
    comps = []
    for f, lead in zip(vars, _COMPONENTS):
        t = _to_tensor(f, dev)
        if lead == (3, 3):
            comps.extend(t[i, j] for i, j in _SYM)
        elif lead == (3,):
            comps.extend(t[i] for i in range(3))
        else:
            comps.append(t)
    return torch.stack(comps).contiguous()


def initialize_tiled_bssn(initial, params, layout):
    """Tile a guard-free global state (``BSSNVariables`` on the active grid,
    or None for flat space) and refresh its guards; returns the tile-major
    device tensor."""
    layout.validate()
    require_bssn_guards(layout, params)
    for axis, names in enumerate(FACE_NAME_PAIRS):
        if any(int(getattr(params, n)) == 1 for n in names):
            if layout.active_shape[axis] == 1:
                raise ValueError("Reduced Cartesian axes must be periodic")
            if min(layout.active_shape[axis], layout.tile_shape[axis] + layout.guard_cells) < 8:
                raise ValueError("Sommerfeld/KO closures require eight accessible inward samples")
    dev = device()
    if initial is None:
        packed = torch.zeros((NCOMP,) + layout.active_shape, dtype=torch.float64, device=dev)
        for c in (0, 3, 5, 6, 17):
            packed[c] = 1.0
    else:
        packed = _pack_global(initial, dev)
        if tuple(packed.shape[1:]) != layout.active_shape:
            raise ValueError(f"Initial state has shape {tuple(packed.shape[1:])}; expected {layout.active_shape}")
    return refresh_tiled_field(tile_global_packed(packed, layout), params, layout)


class TiledEvolver:
    """A tiled state resident in HBM (tile-major packed) advanced with ``step``."""

    def __init__(self, tiled, params, layout, precision="f64"):
        layout.validate()
        require_bssn_guards(layout, params)
        self.layout = layout
        self.params = params
        self.precision = precision_code(precision)
        self.u = tiled
        self.alloc = layout.allocated_shape
        self._cg = _lib.c_grid(self.alloc)
        self._cp = [_lib.c_params(local_parameters(params, layout, idx)) for idx in layout.tiles()]
        self.ustar = torch.empty_like(tiled)
        self.ua = torch.empty_like(tiled)
        self.ub = torch.empty_like(tiled)
        n = max(_lib.work_size(local_parameters(params, layout, (0, 0, 0)), self.alloc, 0), 1)
        self.work = torch.empty(n, dtype=torch.float64, device=tiled.device)

    def refresh(self, buf):
        return refresh_tiled_field(buf, self.params, self.layout)

    def project(self, buf):
        lib = _lib.load()
        st = stream_handle()
        for t in range(self.layout.ntiles):
            _lib.check(lib.bssn_enforce_constraints(ctypes.byref(self._cg), _ptr(buf[t]), _ptr(buf[t]), st))

    def _stage(self, stage, us, out):
        lib = _lib.load()
        st = stream_handle()
        for t in range(self.layout.ntiles):
            _lib.check(lib.bssn_rk4_stage(ctypes.byref(self._cp[t]), ctypes.byref(self._cg), self.precision, stage,
                                          _ptr(us[t]), _ptr(self.u[t]), _ptr(self.ustar[t]), _ptr(out[t]), 0,
                                          _ptr(self.work), st))
        self.refresh(out)

    def step(self, nsteps=1):
        for _ in range(int(nsteps)):
            self.project(self.u)
            self.refresh(self.u)
            self._stage(1, self.u, self.ua)
            self._stage(2, self.ua, self.ub)
            self._stage(3, self.ub, self.ua)
            self._stage(4, self.ua, self.u)
        return self

    def rhs(self, out=None):
        """Tile-local RHS of the refreshed state (compute_tiled_bssn_rhs)."""
        lib = _lib.load()
        st = stream_handle()
        self.refresh(self.u)
        out = torch.empty_like(self.u) if out is None else out
        for t in range(self.layout.ntiles):
            _lib.check(lib.bssn_rhs(ctypes.byref(self._cp[t]), ctypes.byref(self._cg), self.precision,
                                    _ptr(self.u[t]), _ptr(out[t]), _ptr(self.work), st))
        return out

    def constraint_norms(self):
        """Serial-compatible L2 / Linf norms over the owned samples."""
        lib = _lib.load()
        st = stream_handle()
        self.refresh(self.u)
        cf = torch.empty((self.layout.ntiles, 9) + self.alloc, dtype=torch.float64, device=self.u.device)
        for t in range(self.layout.ntiles):
            _lib.check(lib.bssn_constraints(ctypes.byref(self._cp[t]), ctypes.byref(self._cg), _ptr(self.u[t]),
                                            _ptr(cf[t]), st))
        glob = gather_owned(cf, self.layout).contiguous()
        from .constraints import norms_packed
        return dict(zip(NORM_NAMES, norms_packed(glob).cpu().tolist()))

    def global_packed(self):
        return gather_owned(self.u, self.layout)


def tiled_rk4_step(state, params, layout, precision="f64"):
    """Reference evolution.py:123-142 on reference-shaped tiled fields."""
    like = state.conformal_factor
    ev = TiledEvolver(from_reference_tiled(state, layout), params, layout, precision)
    ev.step(1)
    return to_reference_tiled(ev.u, layout, like)


def compute_tiled_bssn_rhs(state, params, layout, precision="f64"):
    """Reference evolution.py:105-108 on reference-shaped tiled fields."""
    like = state.conformal_factor
    ev = TiledEvolver(from_reference_tiled(state, layout), params, layout, precision)
    return to_reference_tiled(ev.rhs(), layout, like)


def tiled_constraint_norms(state, params, layout):
    """Reference evolution.py:173-191."""
    ev = TiledEvolver(from_reference_tiled(state, layout), params, layout)
    return ev.constraint_norms()
