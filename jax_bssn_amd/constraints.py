"""Constraint fields and norms (reference bssn/constraints.py:262-412,
diagnostics/constraints.py:9-36), computed by the HIP constraint kernel and a
two-pass wavefront reduction."""

import ctypes
from typing import NamedTuple

import numpy as np
import torch

from . import _lib
from .time_evolve import _like, _ptr, _same_grid, require_device_f64, require_packed, stream_handle
from .variables import pack


class ConstraintViolations(NamedTuple):
    hamiltonian: object
    momentum: object
    det_gamma: object
    trace_A: object
    gamma_condition: object


NORM_NAMES = ("hamiltonian_l2", "momentum_l2", "det_gamma_l2", "trace_A_l2", "gamma_l2",
              "hamiltonian_linf", "momentum_linf", "det_gamma_linf", "trace_A_linf", "gamma_linf")


def constraints_packed(state, params, out=None, matter=None):
    """(9, nx, ny, nz): H, M_x, M_y, M_z, det-1, tr A, G^x, G^y, G^z."""
    lib = _lib.load()
    require_packed(state)
    if matter is not None:
        _same_grid(matter, state, "matter", 10)
    if out is None:
        out = torch.empty((9,) + tuple(state.shape[1:]), dtype=torch.float64, device=state.device)
    _same_grid(out, state, "out", 9)
    _lib.check(lib.bssn_constraints_with_matter(
        ctypes.byref(_lib.c_params(params)), ctypes.byref(_lib.c_grid(state.shape[1:])), _ptr(state),
        _ptr(matter) if matter is not None else None, _ptr(out), stream_handle()))
    return out


def norms_packed(cfields):
    """Device tensor of the 10 norms of a (9, nx, ny, nz) constraint tensor."""
    lib = _lib.load()
    require_device_f64(cfields, "constraint fields", ncomp=9, ndim=4)
    grid = _lib.c_grid(cfields.shape[1:])
    partial = torch.empty(int(lib.bssn_norm_partial_size(ctypes.byref(grid))), dtype=torch.float64,
                          device=cfields.device)
    norms = torch.empty(10, dtype=torch.float64, device=cfields.device)
    _lib.check(lib.bssn_constraint_norms(ctypes.byref(grid), _ptr(cfields), _ptr(partial), _ptr(norms),
                                         stream_handle()))
    return norms


def _split(cf, like):
    to = (lambda t: t.cpu().numpy()) if isinstance(like, np.ndarray) else (lambda t: t.to(like.device))
    return ConstraintViolations(to(cf[0].clone()), to(cf[1:4].clone()), to(cf[4].clone()),
                                to(cf[5].clone()), to(cf[6:9].clone()))


def compute_all_constraints(vars, params):
    """Reference bssn/constraints.py:387-412."""
    return _split(constraints_packed(pack(vars), params), _like(vars))


def compute_all_constraints_with_matter(vars, params, energy_density, momentum_density):
    """Reference bssn/constraints.py:415-440."""
    from .time_evolve import pack_matter
    state = pack(vars)
    shape = tuple(state.shape[1:])
    zeros = torch.zeros((3, 3) + shape, dtype=torch.float64)
    matter = pack_matter(energy_density, momentum_density, zeros, state.device)
    return _split(constraints_packed(state, params, matter=matter), _like(vars))


def compute_hamiltonian_constraint(vars, params):
    return compute_all_constraints(vars, params).hamiltonian


def compute_constraint_norms(violations):
    """Reference diagnostics/constraints.py:9-36: dict of L2 / Linf norms."""
    from .variables import device
    dev = device()
    h = torch.as_tensor(np.asarray(violations.hamiltonian) if not isinstance(violations.hamiltonian, torch.Tensor)
                        else violations.hamiltonian, dtype=torch.float64).to(dev)
    shape = tuple(h.shape)
    cf = torch.empty((9,) + shape, dtype=torch.float64, device=dev)

    def put(sl, x):
        t = x if isinstance(x, torch.Tensor) else torch.as_tensor(np.asarray(x))
        cf[sl].copy_(t.to(dev, torch.float64).reshape(cf[sl].shape))
    put(0, violations.hamiltonian)
    put(slice(1, 4), violations.momentum)
    put(4, violations.det_gamma)
    put(5, violations.trace_A)
    put(slice(6, 9), violations.gamma_condition)
    vals = norms_packed(cf).cpu().tolist()
    return dict(zip(NORM_NAMES, vals))
