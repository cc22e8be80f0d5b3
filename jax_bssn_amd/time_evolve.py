"""Drop-in RHS / projection / RK4 entry points (reference evolution/time_evolve.py).

Every function takes and returns ``BSSNVariables`` in the reference layout
(NumPy arrays or torch tensors; the result has the kind of the input) and
runs on the GPU through the C ABI of include/bssn_amd.h.  Inputs are never
modified.  ``precision="f32rhs"`` selects the FP32-derivative / FP64-state
variant.  For long runs without host round trips use ``Evolver``, which keeps
the packed state and the RK4 work buffers resident in HBM.
"""

import ctypes

import torch

from . import _lib
from .variables import NCOMP, BSSNParameters, BSSNVariables, device, pack, spatial_shape, unpack

_PRECISIONS = {"f64": _lib.PREC_F64, "f32rhs": _lib.PREC_F32RHS}


def precision_code(precision):
    try:
        return _PRECISIONS[precision]
    except KeyError:
        raise ValueError(f"precision must be one of {sorted(_PRECISIONS)}") from None


def stream_handle():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def require_device_f64(t, name, ncomp=None, ndim=None):
    """Raise ValueError unless ``t`` is a contiguous float64 tensor on a HIP
    device (with ``ndim`` dimensions and ``ncomp`` leading components when
    given): the C ABI takes bare device pointers, so anything else would reach
    the kernels as a host address or a too-small buffer."""
    if not isinstance(t, torch.Tensor):
        raise ValueError(f"{name} must be a torch tensor on the GPU, got {type(t).__name__}")
    if not t.is_cuda:
        raise ValueError(f"{name} must live on a HIP device, not {t.device}")
    if t.dtype != torch.float64:
        raise ValueError(f"{name} must be float64, got {t.dtype}")
    if not t.is_contiguous():
        raise ValueError(f"{name} must be contiguous")
    if ndim is not None and t.dim() != ndim:
        raise ValueError(f"{name} must have {ndim} dimensions, got shape {tuple(t.shape)}")
    if ncomp is not None and (t.dim() < 1 or t.shape[0] != ncomp):
        raise ValueError(f"{name} must have {ncomp} leading components, got shape {tuple(t.shape)}")
    return t


def require_packed(t, name="state"):
    """A packed (21, nx, ny, nz) FP64 device state (see require_device_f64)."""
    return require_device_f64(t, name, ncomp=NCOMP, ndim=4)


def _same_grid(t, state, name, ncomp):
    require_device_f64(t, name, ncomp=ncomp, ndim=4)
    if tuple(t.shape[1:]) != tuple(state.shape[1:]) or t.device != state.device:
        raise ValueError(f"{name} has shape {tuple(t.shape)} on {t.device}; expected "
                         f"({ncomp},) + {tuple(state.shape[1:])} on {state.device}")


def _like(v):
    return v.conformal_factor


def _scratch(params, shape, what, device):
    n = _lib.work_size(params, shape, what)
    return torch.empty(max(n, 1), dtype=torch.float64, device=device)


def rhs_packed(state, params, precision="f64", out=None):
    """RHS of a packed (21, nx, ny, nz) device state."""
    lib = _lib.load()
    require_packed(state)
    shape = tuple(state.shape[1:])
    if out is None:
        out = torch.empty_like(state)
    _same_grid(out, state, "out", NCOMP)
    work = _scratch(params, shape, 0, state.device)
    _lib.check(lib.bssn_rhs(ctypes.byref(_lib.c_params(params)), ctypes.byref(_lib.c_grid(shape)),
                            precision_code(precision), _ptr(state), _ptr(out), _ptr(work), stream_handle()))
    return out


def pack_matter(energy_density, momentum_density, spatial_stress, device_):
    """(rho[n], S_i[3,n], S_ij[3,3,n]) -> (10, nx, ny, nz) device tensor."""
    def t(x):
        x = x if isinstance(x, torch.Tensor) else torch.as_tensor(__import__("numpy").asarray(x))
        return x.to(device=device_, dtype=torch.float64)
    rho = t(energy_density)
    m = torch.empty((10,) + tuple(rho.shape), dtype=torch.float64, device=device_)
    m[0].copy_(rho)
    m[1:4].copy_(t(momentum_density))
    Sij = t(spatial_stress)
    for c, (i, j) in enumerate(((0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2))):
        m[4 + c].copy_(Sij[i, j])
    return m


def rhs_with_matter_packed(state, matter, params, precision="f64", out=None):
    lib = _lib.load()
    require_packed(state)
    _same_grid(matter, state, "matter", 10)
    shape = tuple(state.shape[1:])
    if out is None:
        out = torch.empty_like(state)
    _same_grid(out, state, "out", NCOMP)
    work = _scratch(params, shape, 0, state.device)
    _lib.check(lib.bssn_rhs_with_matter(ctypes.byref(_lib.c_params(params)), ctypes.byref(_lib.c_grid(shape)),
                                        precision_code(precision), _ptr(state), _ptr(matter), _ptr(out),
                                        _ptr(work), stream_handle()))
    return out


def project_packed(state, out=None):
    lib = _lib.load()
    require_packed(state)
    if out is None:
        out = torch.empty_like(state)
    _same_grid(out, state, "out", NCOMP)
    _lib.check(lib.bssn_enforce_constraints(ctypes.byref(_lib.c_grid(state.shape[1:])), _ptr(state),
                                            _ptr(out), stream_handle()))
    return out


def evolve_packed(state, params, nsteps, work=None, precision="f64"):
    """rk4_step applied nsteps times to a packed device state, in place."""
    lib = _lib.load()
    require_packed(state)
    need = _lib.work_size(params, tuple(state.shape[1:]), 1)
    if work is None:
        work = _scratch(params, tuple(state.shape[1:]), 1, state.device)
    require_device_f64(work, "work")
    if work.numel() < need or work.device != state.device:
        raise ValueError(f"work holds {work.numel()} doubles on {work.device}; bssn_evolve needs {need} "
                         f"on {state.device}")
    _lib.check(lib.bssn_evolve(ctypes.byref(_lib.c_params(params)), ctypes.byref(_lib.c_grid(state.shape[1:])),
                               precision_code(precision), int(nsteps), _ptr(state), _ptr(work),
                               stream_handle()))
    return state


def compute_bssn_rhs(vars, params, precision="f64"):
    """Reference time_evolve.py:41-57 (vacuum RHS + Sommerfeld faces)."""
    state = pack(vars)
    return unpack(rhs_packed(state, params, precision), like=_like(vars))


def compute_bssn_rhs_with_matter(vars, params, energy_density, momentum_density, spatial_stress,
                                 precision="f64"):
    """Reference time_evolve.py:60-91 (Eulerian matter sources rho, S_i, S_ij)."""
    state = pack(vars)
    matter = pack_matter(energy_density, momentum_density, spatial_stress, state.device)
    return unpack(rhs_with_matter_packed(state, matter, params, precision), like=_like(vars))


def enforce_algebraic_constraints(vars):
    """Reference time_evolve.py:134-141 (det = 1, then trace-free A)."""
    state = pack(vars)
    return unpack(project_packed(state, state), like=_like(vars))


def rk4_step(vars, params, precision="f64"):
    """Reference time_evolve.py:144-217."""
    state = pack(vars)
    evolve_packed(state, params, 1, precision=precision)
    return unpack(state, like=_like(vars))


def evolve(vars, params, nsteps, precision="f64"):
    """``nsteps`` applications of rk4_step, with the state resident in HBM."""
    state = pack(vars)
    evolve_packed(state, params, nsteps, precision=precision)
    return unpack(state, like=_like(vars))


class Evolver:
    """Packed device state plus RK4 work buffers for repeated stepping.

    ``stage(s)`` exposes the single fused RK4 stage (bssn_rk4_stage).
    """

    def __init__(self, vars_or_state, params, precision="f64"):
        self.params = params
        self.precision = precision
        if isinstance(vars_or_state, torch.Tensor):
            self.state = require_packed(vars_or_state)
        else:
            self.state = pack(vars_or_state)
        self.shape = tuple(self.state.shape[1:])
        n = _lib.work_size(params, self.shape, 1)
        self.workbuf = torch.empty(n, dtype=torch.float64, device=self.state.device)
        # the three RK4 state buffers, then the stage scratch
        self.work = self.workbuf[:3 * NCOMP * int(torch.Size(self.shape).numel())].view((3, NCOMP) + self.shape)
        self.stage_scratch = self.workbuf[3 * NCOMP * int(torch.Size(self.shape).numel()):]
        self._cp = _lib.c_params(params)
        self._cg = _lib.c_grid(self.shape)
        self._prec = precision_code(precision)

    def step(self, nsteps=1):
        lib = _lib.load()
        _lib.check(lib.bssn_evolve(ctypes.byref(self._cp), ctypes.byref(self._cg), self._prec, int(nsteps),
                                   _ptr(self.state), _ptr(self.workbuf), stream_handle()))
        return self

    def stage(self, s, u_s, u0, ustar, u_next, reproject=False):
        lib = _lib.load()
        for name, t in (("u_s", u_s), ("u0", u0), ("ustar", ustar), ("u_next", u_next)):
            if t is not None:
                _same_grid(t, self.state, name, NCOMP)
        _lib.check(lib.bssn_rk4_stage(ctypes.byref(self._cp), ctypes.byref(self._cg), self._prec, int(s),
                                      _ptr(u_s), _ptr(u0) if u0 is not None else None, _ptr(ustar),
                                      _ptr(u_next), 1 if reproject else 0, _ptr(self.stage_scratch),
                                      stream_handle()))

    def variables(self, like=None):
        return unpack(self.state, like=like)
