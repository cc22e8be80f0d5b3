"""Forward-mode differentiation of the evolution (exact tangent-linear maps
through the HIP kernels) and a torch.autograd.Function for gradients of the
L2 Hamiltonian constraint with respect to initial-data parameters.

The reference differentiates its JAX evolution with autodiff.  Here the
per-point RHS, the RK4 stage update and the projection are instantiated with
dual numbers (value and derivatives along one or two directions), so one
tangent run gives the exact derivatives of the whole evolution along two
initial-data directions at once; p parameters take ceil(p / 2) tangent runs
and no checkpoints (nothing is replayed backwards).  kappa = 0 and vacuum
only.

A tangent argument ``dstate`` is either one direction (the shape of the
state) or a stack of up to two directions (``(ndir,) + state.shape``);
results keep the form they were given in.
"""

import ctypes
import math

import numpy as np
import torch

from . import _lib
from .time_evolve import _ptr, require_packed, stream_handle
from .variables import NCOMP, device


MAX_DIRS = 2


def _prep(t):
    return t.to(device(), torch.float64).contiguous().clone()


def _dirs(state, dstate):
    """(stacked (ndir, 21, ...) tangent tensor, whether the input was one
    direction)."""
    single = dstate.dim() == state.dim()
    d = _prep(dstate.unsqueeze(0) if single else dstate)
    if tuple(d.shape[1:]) != tuple(state.shape) or not 1 <= d.shape[0] <= MAX_DIRS:
        raise ValueError(f"tangent shape {tuple(dstate.shape)} does not fit state {tuple(state.shape)} "
                         f"with 1..{MAX_DIRS} directions")
    return d, single


def rhs_tangent(state, dstate, params):
    """(RHS(state), dRHS(state)[dstate]) of packed device states."""
    state = _prep(state)
    d, single = _dirs(state, dstate)
    rhs, drhs = torch.empty_like(state), torch.empty_like(d)
    _lib.check(_lib.load().bssn_rhs_tangent(ctypes.byref(_lib.c_params(params)),
                                            ctypes.byref(_lib.c_grid(state.shape[1:])), int(d.shape[0]),
                                            _ptr(state), _ptr(d), _ptr(rhs), _ptr(drhs), stream_handle()))
    return rhs, (drhs[0] if single else drhs)


def evolve_tangent(state, dstate, params, nsteps):
    """``nsteps`` RK4 steps of a packed state with its tangents carried along."""
    state = _prep(state)
    d, single = _dirs(state, dstate)
    lib = _lib.load()
    grid = _lib.c_grid(state.shape[1:])
    size = int(lib.bssn_tangent_work_size(ctypes.byref(grid), int(d.shape[0])))
    if size < 0:
        _lib.check(size)
    work = torch.empty(size, dtype=torch.float64, device=state.device)
    _lib.check(lib.bssn_evolve_tangent(ctypes.byref(_lib.c_params(params)), ctypes.byref(grid), int(nsteps),
                                       int(d.shape[0]), _ptr(state), _ptr(d), _ptr(work), stream_handle()))
    return state, (d[0] if single else d)


def hamiltonian_l2_tangent(state, dstate, params):
    """(L, dL) with L = sqrt(mean H^2) over all points and dL its derivative
    along each direction of dstate (mean(H dH) / L); dL is a float for one
    direction and a list for a stack."""
    require_packed(state)
    d, single = _dirs(state, dstate)
    shape = tuple(state.shape[1:])
    H = torch.empty(shape, dtype=torch.float64, device=state.device)
    dH = torch.empty((d.shape[0],) + shape, dtype=torch.float64, device=state.device)
    _lib.check(_lib.load().bssn_hamiltonian_tangent(ctypes.byref(_lib.c_params(params)),
                                                    ctypes.byref(_lib.c_grid(shape)), int(d.shape[0]),
                                                    _ptr(state), _ptr(d), _ptr(H), _ptr(dH), stream_handle()))
    L = torch.sqrt(torch.mean(H * H))
    dL = [float(torch.mean(H * dH[l]) / L) for l in range(d.shape[0])]
    return float(L), (dL[0] if single else dL)


class HamiltonianL2AfterSteps(torch.autograd.Function):
    """L2 Hamiltonian constraint after ``nsteps`` RK4 steps of the initial
    data ``initial(theta)``, differentiable in theta.

    ``initial(theta)`` (theta a NumPy vector) returns the packed initial
    state and, for every component of theta, the packed derivative of the
    initial state with respect to it.  The forward pass runs one tangent
    evolution per pair of components; backward multiplies the stored
    gradient by the incoming gradient.
    """

    @staticmethod
    def forward(ctx, theta, initial, params, nsteps):
        u0, tangents = initial(theta.detach().cpu().numpy())
        value, grads = None, []
        for k in range(0, len(tangents), MAX_DIRS):
            du0 = torch.stack(tangents[k:k + MAX_DIRS])
            u, du = evolve_tangent(u0, du0, params, nsteps)
            value, dL = hamiltonian_l2_tangent(u, du, params)
            grads.extend(dL)
        ctx.save_for_backward(torch.tensor(grads, dtype=theta.dtype))
        return torch.tensor(value, dtype=theta.dtype)

    @staticmethod
    def backward(ctx, grad_out):
        (grads,) = ctx.saved_tensors
        return grad_out * grads, None, None, None


def puncture_initial(X, Y, Z, centre=(0.0, 0.0, 0.0), width=1.0, lapse_centre=(0.0, 0.0, 0.0)):
    """``initial(theta)`` for theta = (M, eps): one zero-momentum puncture of
    mass M at ``centre`` (W = psi^-2, alpha = W) plus eps exp(-|x - c|^2 /
    width^2) on the lapse, with the analytic derivatives
    dW/dM = -psi^-3 / r and dalpha/deps = the Gaussian."""
    X, Y, Z = (np.asarray(a, dtype=np.float64) for a in (X, Y, Z))
    r = np.sqrt((X - centre[0]) ** 2 + (Y - centre[1]) ** 2 + (Z - centre[2]) ** 2)
    gauss = np.exp(-((X - lapse_centre[0]) ** 2 + (Y - lapse_centre[1]) ** 2 + (Z - lapse_centre[2]) ** 2)
                   / width ** 2)

    def initial(theta):
        M, eps = float(theta[0]), float(theta[1])
        psi = 1.0 + M / (2.0 * r)
        W = psi ** -2
        dev = device()
        u = torch.zeros((NCOMP,) + X.shape, dtype=torch.float64, device=dev)
        for c in (0, 3, 5):
            u[c] = 1.0
        u[6] = torch.as_tensor(W, device=dev)
        u[17] = torch.as_tensor(W + eps * gauss, device=dev)
        dM = torch.zeros_like(u)
        dWdM = torch.as_tensor(-psi ** -3 / r, device=dev)
        dM[6] = dWdM
        dM[17] = dWdM
        de = torch.zeros_like(u)
        de[17] = torch.as_tensor(gauss, device=dev)
        return u, [dM, de]
    return initial
