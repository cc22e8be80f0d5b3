"""Axisymmetric (z-axis) Cartoon evolution on the GPU (reference
cartoon/axisymmetry/).

The evolved state is the compact ``y = 0`` half-plane of the reference
(reconstruction.py:1-145): extents ``(nx, 1, nz)``, x indices 0..3 parity
ghosts and 4.. the positive radial half-cell centres.  Each RHS evaluation
reconstructs the nine y planes the Cartesian stencils need
(``bssn_cartoon_axi_reconstruct``), evaluates the unchanged BSSN RHS on the
centre plane only (``bssn_rhs_region``; the reference evaluates all nine and
keeps the centre, evolution.py:18-28), and projects it back with parity
ghosts (``bssn_cartoon_axi_project``).  The RK4 step follows evolution.py:
36-65 (projection and ghost refill at every stage input).
"""

import ctypes
import math

import numpy as np
import torch

from . import _lib
from .time_evolve import _ptr, precision_code, stream_handle
from .variables import BSSNVariables, device, pack, unpack

AXISYMMETRIC_GHOST_CELLS = 4
AXISYMMETRIC_SUPPORT_SIZE = 9
AXISYMMETRIC_CENTER = 4
AXISYMMETRIC_OUTER_BUFFER_CELLS = 3
VECTOR_RADIAL_REFLECTION_PARITY = (-1.0, -1.0, 1.0)


def validate_axisymmetric_grid(vars, params):
    """Reference reconstruction.py:73-113 (ValueError on the same inputs)."""
    shape = tuple(np.shape(vars.conformal_factor))
    nx, ny, nz = shape
    if ny != 1:
        raise ValueError("axisymmetric Cartoon state requires a singleton y plane")
    if nx - AXISYMMETRIC_GHOST_CELLS < 6:
        raise ValueError("axisymmetric Cartoon mode requires at least six radial points")
    if nz < 9:
        raise ValueError("axisymmetric Cartoon mode requires at least nine z points")
    expected = ((vars.conformal_metric, (3, 3) + shape, "conformal_metric"),
                (vars.conformal_factor, shape, "conformal_factor"),
                (vars.traceless_K, (3, 3) + shape, "traceless_K"), (vars.trace_K, shape, "trace_K"),
                (vars.conformal_connection, (3,) + shape, "conformal_connection"), (vars.lapse, shape, "lapse"),
                (vars.shift, (3,) + shape, "shift"))
    for field, want, name in expected:
        if tuple(np.shape(field)) != want:
            raise ValueError(f"{name} has shape {tuple(np.shape(field))}; expected {want}")
    if params.xl_bc != 0:
        raise ValueError("axisymmetric x-left is a parity ghost region")
    if params.xr_bc != 1:
        raise ValueError("axisymmetric x-right must use the Sommerfeld boundary")
    if params.yl_bc != 0 or params.yr_bc != 0:
        raise ValueError("axisymmetric reconstructed y faces must be periodic")
    if params.zl_bc not in (0, 1) or params.zr_bc not in (0, 1):
        raise ValueError("axisymmetric z faces must be periodic or Sommerfeld")
    if params.mad_q != 1:
        raise ValueError("axisymmetric Cartoon requires mad_q=1")
    dx = float(params.dx)
    if not math.isfinite(dx) or dx <= 0.0:
        raise ValueError("axisymmetric Cartoon requires finite dx > 0")
    if not math.isclose(params.x_min, -3.5 * dx, rel_tol=1e-5, abs_tol=1e-8):
        raise ValueError("axisymmetric Cartoon requires x_min=-3.5*dx")
    if not math.isclose(params.y_min, -4.0 * dx, rel_tol=1e-5, abs_tol=1e-8):
        raise ValueError("axisymmetric Cartoon requires y_min=-4*dx")


def _support_shape(compact_shape):
    nx, _, nz = compact_shape
    return (nx, AXISYMMETRIC_SUPPORT_SIZE, nz)


class AxisymmetricEvolver:
    """A compact axisymmetric state resident in HBM, advanced with ``step``."""

    def __init__(self, state, params, precision="f64", validate=True):
        dev = device()
        if isinstance(state, BSSNVariables):
            if validate:
                validate_axisymmetric_grid(state, params)
            state = pack(state)
        self.u = state.to(dev, torch.float64).contiguous().clone()
        self.shape = tuple(self.u.shape[1:])
        self.params = params
        self.precision = precision_code(precision)
        self._cg = _lib.c_grid(self.shape)
        sshape = _support_shape(self.shape)
        self._cs = _lib.c_grid(sshape)
        self._cp = _lib.c_params(params)
        self.ustar = torch.empty_like(self.u)
        self.ua = torch.empty_like(self.u)
        self.ub = torch.empty_like(self.u)
        self.k = torch.empty_like(self.u)
        self.sup = torch.empty((21,) + sshape, dtype=torch.float64, device=dev)
        self.sup_rhs = torch.empty_like(self.sup)
        self.work = torch.empty(max(_lib.work_size(params, sshape, 0), 1), dtype=torch.float64, device=dev)
        self._lo = (ctypes.c_int32 * 3)(AXISYMMETRIC_GHOST_CELLS, AXISYMMETRIC_CENTER, 0)
        self._hi = (ctypes.c_int32 * 3)(self.shape[0], AXISYMMETRIC_CENTER + 1, self.shape[2])

    def fill(self, buf):
        _lib.check(_lib.load().bssn_cartoon_axi_fill_ghosts(ctypes.byref(self._cg), _ptr(buf), stream_handle()))

    def rhs(self, us, out):
        """Compact RHS of the compact state ``us`` into ``out``."""
        lib = _lib.load()
        st = stream_handle()
        _lib.check(lib.bssn_cartoon_axi_reconstruct(ctypes.byref(self._cg), float(self.params.dx),
                                                    float(self.params.z_min), _ptr(us), _ptr(self.sup), st))
        _lib.check(lib.bssn_rhs_region(ctypes.byref(self._cp), ctypes.byref(self._cs), self.precision, self._lo,
                                       self._hi, _ptr(self.sup), _ptr(self.sup_rhs), _ptr(self.work), st))
        _lib.check(lib.bssn_cartoon_axi_project(ctypes.byref(self._cg), _ptr(self.sup_rhs), _ptr(out), st))
        return out

    def _combine(self, stage, us, out):
        self.rhs(us, self.k)
        _lib.check(_lib.load().bssn_rk4_combine(ctypes.byref(self._cg), float(self.params.dt), stage, _ptr(self.k),
                                                _ptr(self.u), _ptr(self.ustar), _ptr(out), 0, stream_handle()))
        self.fill(out)

    def step(self, nsteps=1):
        lib = _lib.load()
        for _ in range(int(nsteps)):
            _lib.check(lib.bssn_enforce_constraints(ctypes.byref(self._cg), _ptr(self.u), _ptr(self.u),
                                                    stream_handle()))
            self.fill(self.u)
            self._combine(1, self.u, self.ua)
            self._combine(2, self.ua, self.ub)
            self._combine(3, self.ub, self.ua)
            self._combine(4, self.ua, self.u)
        return self

    def variables(self, like=None):
        return unpack(self.u, like)


def _like(vars):
    return vars.conformal_factor


def fill_axisymmetric_ghosts(vars):
    """Reference reconstruction.py:131-145."""
    s = pack(vars).contiguous()
    _lib.check(_lib.load().bssn_cartoon_axi_fill_ghosts(ctypes.byref(_lib.c_grid(s.shape[1:])), _ptr(s),
                                                        stream_handle()))
    return unpack(s, _like(vars))


def reconstruct_axisymmetric_support(vars, params):
    """Reference reconstruction.py:255-292: the nine-plane support state."""
    s = pack(vars).contiguous()
    sup = torch.empty((21,) + _support_shape(tuple(s.shape[1:])), dtype=torch.float64, device=s.device)
    _lib.check(_lib.load().bssn_cartoon_axi_reconstruct(ctypes.byref(_lib.c_grid(s.shape[1:])), float(params.dx),
                                                        float(params.z_min), _ptr(s), _ptr(sup), stream_handle()))
    return unpack(sup, _like(vars))


def project_axisymmetric_rhs(rhs):
    """Reference reconstruction.py:295-304: compact RHS from a support RHS."""
    s = pack(rhs).contiguous()
    nx, _, nz = s.shape[1:]
    out = torch.empty((21, nx, 1, nz), dtype=torch.float64, device=s.device)
    _lib.check(_lib.load().bssn_cartoon_axi_project(ctypes.byref(_lib.c_grid((nx, 1, nz))), _ptr(s), _ptr(out),
                                                    stream_handle()))
    return unpack(out, _like(rhs))


def compute_axisymmetric_rhs(vars, params, precision="f64"):
    """Reference axisymmetry/evolution.py:18-28."""
    validate_axisymmetric_grid(vars, params)
    ev = AxisymmetricEvolver(vars, params, precision=precision, validate=False)
    return unpack(ev.rhs(ev.u, ev.k), _like(vars))


def axisymmetric_rk4_step(vars, params, precision="f64"):
    """Reference axisymmetry/evolution.py:36-65."""
    ev = AxisymmetricEvolver(vars, params, precision=precision)
    ev.step(1)
    return ev.variables(_like(vars))


def _map_fields(vars, fn):
    par = np.asarray(VECTOR_RADIAL_REFLECTION_PARITY)
    out = []
    for f in vars:
        a = f.cpu().numpy() if isinstance(f, torch.Tensor) else np.asarray(f)
        rank = a.ndim - 3
        p = np.asarray(1.0) if rank == 0 else (par if rank == 1 else par[:, None] * par[None, :])
        out.append(fn(a, p.reshape(p.shape + (1, 1)) if rank else p))
    return BSSNVariables(*out)


def compact_axisymmetric_state(vars):
    """Reference reconstruction.py:116-128: a complete signed x-z plane (even
    x extent, y = 1) into compact storage (data layout only)."""
    nx, ny, _ = np.shape(vars.conformal_factor)
    if ny != 1 or nx % 2:
        raise ValueError("input must be an even-length signed x plane with y=1")

    def fn(a, p):
        pos = a[..., nx // 2:, 0, :]
        ghosts = np.flip(pos[..., :AXISYMMETRIC_GHOST_CELLS, :], axis=-2) * p
        return np.concatenate((ghosts, pos), axis=-2)[..., :, None, :]
    return _map_fields(vars, fn)


def expand_axisymmetric_plane(vars):
    """Reference reconstruction.py:307-314: compact fields onto the complete
    signed x-z plane (data layout only)."""
    def fn(a, p):
        pos = a[..., AXISYMMETRIC_GHOST_CELLS:, 0, :]
        return np.concatenate((np.flip(pos, axis=-2) * p, pos), axis=-2)[..., :, None, :]
    return _map_fields(vars, fn)
