"""Stage-synchronous fixed mesh refinement on one GPU (reference fmr/).

Names and meaning follow the reference: ``FMRPatchSpec`` / ``FMRHierarchySpec``
(fmr/refinement.py:15-31), ``hierarchy_parents`` (:34-51),
``fine_active_shape`` (:70-74), ``evolution_parameters`` (fmr/tiled.py:39-48)
and the step ``fmr_rk4_step`` (fmr/tiled.py:182-209, the frozen serial
statement tests/support/fmr_reference.py:256-310).  Level storage is that of
the serial statement: the root is the whole (unpadded) grid, every child a
vertex-centred 2:1 patch of ``2 (hi - lo) + 1`` active vertices per axis plus
``ghost_width`` guard points on each side.

All arithmetic runs in HIP kernels through the C ABI: the fused RK4 stage
kernels per level (children evolve the bulk equations with periodic face
codes inside their allocation, the MAD weight ``(dx_finest / dx)^4`` on
coarser levels), ``bssn_fmr_prolong`` for the child guard shells (root to
leaf after every stage) and ``bssn_fmr_restrict`` for the injection (leaf
to root after the step).
"""

import ctypes
import math
from typing import NamedTuple

import torch

from . import _lib
from .time_evolve import _ptr, precision_code, stream_handle
from .variables import BSSNVariables, device, pack, unpack

FACE_NAMES = ("xl_bc", "xr_bc", "yl_bc", "yr_bc", "zl_bc", "zr_bc")


class FMRPatchSpec(NamedTuple):
    """Inclusive parent-active-index bounds of one child patch."""

    coarse_lo: tuple
    coarse_hi: tuple
    refinement_ratio: int = 2
    ghost_width: int = 4
    parent: object = None


class FMRHierarchySpec(NamedTuple):
    """Parent-before-child patches and the hierarchy-wide MAD choice."""

    patches: tuple
    use_mad: bool = True


def _validate_spec(spec):
    if len(spec.coarse_lo) != 3 or len(spec.coarse_hi) != 3:
        raise ValueError("FMR patch bounds must be three-dimensional")
    if spec.refinement_ratio != 2:
        raise ValueError("FMR refinement ratio must be 2:1")
    if spec.ghost_width != 4:
        raise ValueError("the BSSN FMR configuration requires four guard cells")
    if any(h < l for l, h in zip(spec.coarse_lo, spec.coarse_hi)):
        raise ValueError("coarse_hi must not precede coarse_lo")


def fine_active_shape(spec):
    _validate_spec(spec)
    return tuple(spec.refinement_ratio * (h - l) + 1 for l, h in zip(spec.coarse_lo, spec.coarse_hi))


def fine_allocation_shape(spec):
    return tuple(n + 2 * spec.ghost_width for n in fine_active_shape(spec))


def fine_active_slice(spec):
    g = spec.ghost_width
    return tuple(slice(g, g + n) for n in fine_active_shape(spec))


def hierarchy_parents(hierarchy):
    """Flat parent index per patch (root: -1); omitted parents form a chain."""
    parents = [-1]
    for child, spec in enumerate(hierarchy.patches, start=1):
        parent = child - 1 if spec.parent is None else spec.parent
        if not isinstance(parent, int) or not 0 <= parent < child:
            raise ValueError("FMR patches must follow their parent in the hierarchy")
        parents.append(parent)
    for child, spec in enumerate(hierarchy.patches, start=1):
        for other in range(1, child):
            sib = hierarchy.patches[other - 1]
            if parents[child] == parents[other] and all(
                    max(a, b) <= min(c, d) for a, b, c, d in zip(spec.coarse_lo, sib.coarse_lo,
                                                                  spec.coarse_hi, sib.coarse_hi)):
                raise ValueError("sibling FMR patches must not share owned vertices")
    return tuple(parents)


def evolution_parameters(parameters, hierarchy):
    """MAD weight from the finest spacing on every level; periodic child faces."""
    finest = min(p.dx for p in parameters)
    out = []
    for level, p in enumerate(parameters):
        changes = dict(mad_q=(finest / p.dx) ** 4 if hierarchy.use_mad else 1.0)
        if level:
            changes.update({n: 0 for n in FACE_NAMES})
        out.append(p._replace(**changes))
    return tuple(out)


def validate_hierarchy(shapes, parameters, hierarchy):
    """Reference-style checks (fmr/tiled.py:51-104, fmr_reference.py:188-253)."""
    parents = hierarchy_parents(hierarchy)
    if len(parameters) != len(parents) or len(shapes) != len(parents):
        raise ValueError("FMR requires one state and one parameter tuple per patch")
    dt = float(parameters[0].dt)
    if not math.isfinite(dt) or dt <= 0:
        raise ValueError("the shared FMR timestep must be finite and positive")
    for level, (shape, p) in enumerate(zip(shapes, parameters)):
        if not math.isclose(float(p.dt), dt, rel_tol=1e-5, abs_tol=1e-8):
            raise ValueError("stage-synchronous FMR requires identical timesteps")
        if not level:
            continue
        spec = hierarchy.patches[level - 1]
        if tuple(shape) != fine_allocation_shape(spec):
            raise ValueError(f"level {level} allocation shape {tuple(shape)} does not match "
                             f"expected {fine_allocation_shape(spec)}")
        parent = parents[level]
        if not math.isclose(float(p.dx) * 2, float(parameters[parent].dx), rel_tol=1e-5, abs_tol=1e-8):
            raise ValueError("successive FMR spacings must differ by exactly two")
        pshape = shapes[parent] if parent == 0 else fine_active_shape(hierarchy.patches[parent - 1])
        if any(lo <= 0 or hi >= n - 1 for lo, hi, n in zip(spec.coarse_lo, spec.coarse_hi, pshape)):
            raise ValueError("every child patch must be strictly inside its parent")


class FMREvolver:
    """A hierarchy resident in HBM, advanced with ``step``.

    ``states``: one entry per patch (``BSSNVariables`` in the reference
    layout, or packed (21, nx, ny, nz) device tensors), root first, children
    padded by ``ghost_width``.  ``prepare=True`` fills the child guards from
    their parents once (the reference's initialisation refresh, without
    projection); ``fmr_rk4_step`` starts from the states as given.
    """

    def __init__(self, states, parameters, hierarchy, precision="f64", prepare=True):
        self.hierarchy = hierarchy
        self.parents = hierarchy_parents(hierarchy)
        self.precision = precision_code(precision)
        dev = device()
        packed = [s if isinstance(s, torch.Tensor) and s.dim() == 4 else pack(s) for s in states]
        self.u = [s.to(dev, torch.float64).contiguous().clone() for s in packed]
        shapes = [tuple(s.shape[1:]) for s in self.u]
        validate_hierarchy(shapes, parameters, hierarchy)
        self.params = evolution_parameters(tuple(parameters), hierarchy)
        self._cp = [_lib.c_params(p) for p in self.params]
        self._cg = [_lib.c_grid(s) for s in shapes]
        self.ustar = [torch.empty_like(s) for s in self.u]
        self.ua = [torch.empty_like(s) for s in self.u]
        self.ub = [torch.empty_like(s) for s in self.u]
        self.work = [torch.empty(max(_lib.work_size(p, s, 0), 1), dtype=torch.float64, device=dev)
                     for p, s in zip(self.params, shapes)]
        self._lo = [None] + [(ctypes.c_int32 * 3)(*spec.coarse_lo) for spec in hierarchy.patches]
        self._periodic = (ctypes.c_int32 * 3)(*[int(self.params[0].xl_bc == 0 and self.params[0].xr_bc == 0),
                                                int(self.params[0].yl_bc == 0 and self.params[0].yr_bc == 0),
                                                int(self.params[0].zl_bc == 0 and self.params[0].zr_bc == 0)])
        self._none = (ctypes.c_int32 * 3)(0, 0, 0)
        if prepare:
            self.fill(self.u)

    def _ghost(self, level):
        return 0 if level == 0 else self.hierarchy.patches[level - 1].ghost_width

    def fill(self, bufs):
        """Child guard shells from their parents, root to leaf."""
        lib = _lib.load()
        st = stream_handle()
        for child in range(1, len(bufs)):
            par = self.parents[child]
            _lib.check(lib.bssn_fmr_prolong(
                ctypes.byref(self._cg[par]), self._ghost(par), self._periodic if par == 0 else self._none,
                _ptr(bufs[par]), ctypes.byref(self._cg[child]), self._lo[child],
                self.hierarchy.patches[child - 1].ghost_width, _ptr(bufs[child]), st))

    def project(self, bufs):
        lib = _lib.load()
        st = stream_handle()
        for b, g in zip(bufs, self._cg):
            _lib.check(lib.bssn_enforce_constraints(ctypes.byref(g), _ptr(b), _ptr(b), st))

    def _stage(self, stage, us, out):
        lib = _lib.load()
        st = stream_handle()
        for lv in range(len(self.u)):
            _lib.check(lib.bssn_rk4_stage(
                ctypes.byref(self._cp[lv]), ctypes.byref(self._cg[lv]), self.precision, stage,
                _ptr(us[lv]), _ptr(self.u[lv]), _ptr(self.ustar[lv]), _ptr(out[lv]), 0,
                _ptr(self.work[lv]), st))
        self.fill(out)

    def step(self, nsteps=1):
        lib = _lib.load()
        st = stream_handle()
        for _ in range(int(nsteps)):
            self.project(self.u)
            self.fill(self.u)
            self._stage(1, self.u, self.ua)
            self._stage(2, self.ua, self.ub)
            self._stage(3, self.ub, self.ua)
            # stage 4 reads ustar and writes the result over u (u0 is unused)
            for lv in range(len(self.u)):
                _lib.check(lib.bssn_rk4_stage(
                    ctypes.byref(self._cp[lv]), ctypes.byref(self._cg[lv]), self.precision, 4,
                    _ptr(self.ua[lv]), _ptr(self.u[lv]), _ptr(self.ustar[lv]), _ptr(self.u[lv]), 0,
                    _ptr(self.work[lv]), st))
            for child in range(len(self.u) - 1, 0, -1):
                par = self.parents[child]
                _lib.check(lib.bssn_fmr_restrict(
                    ctypes.byref(self._cg[child]), self._lo[child], self.hierarchy.patches[child - 1].ghost_width,
                    _ptr(self.u[child]), ctypes.byref(self._cg[par]), self._ghost(par), _ptr(self.u[par]), st))
                _lib.check(lib.bssn_enforce_constraints(ctypes.byref(self._cg[par]), _ptr(self.u[par]),
                                                        _ptr(self.u[par]), st))
            self.fill(self.u)
        return self

    def variables(self, like=None):
        return [unpack(s, like) for s in self.u]


def fmr_rk4_step(states, parameters, hierarchy, precision="f64"):
    """Advance every patch through one synchronized RK4 step; returns a list
    of ``BSSNVariables`` (NumPy when the inputs are NumPy)."""
    like = states[0].conformal_factor if isinstance(states[0], BSSNVariables) else None
    ev = FMREvolver(states, parameters, hierarchy, precision=precision, prepare=False)
    ev.step(1)
    return ev.variables(like)
