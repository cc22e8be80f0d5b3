"""Forward-mode differentiation of the evolution (exact tangent-linear maps
through the HIP kernels) and a torch.autograd.Function for gradients of the
L2 Hamiltonian constraint with respect to initial-data parameters.

The reference differentiates its JAX evolution with autodiff.  Here the
per-point RHS, the RK4 stage update and the projection are instantiated with
dual numbers (value, derivative along one direction), so one tangent run
gives the exact derivative of the whole evolution along one initial-data
direction; p parameters take p tangent runs and no checkpoints (nothing is
replayed backwards).  kappa = 0 and vacuum only.
"""

import ctypes
import math

import numpy as np
import torch

from . import _lib
from .time_evolve import _ptr, stream_handle
from .variables import NCOMP, device


def _prep(t):
    return t.to(device(), torch.float64).contiguous().clone()


def rhs_tangent(state, dstate, params):
    """(RHS(state), dRHS(state)[dstate]) of packed device states."""
    state, dstate = _prep(state), _prep(dstate)
    rhs, drhs = torch.empty_like(state), torch.empty_like(state)
    _lib.check(_lib.load().bssn_rhs_tangent(ctypes.byref(_lib.c_params(params)),
                                            ctypes.byref(_lib.c_grid(state.shape[1:])), _ptr(state), _ptr(dstate),
                                            _ptr(rhs), _ptr(drhs), stream_handle()))
    return rhs, drhs


def evolve_tangent(state, dstate, params, nsteps):
    """``nsteps`` RK4 steps of a packed state with its tangent carried along."""
    state, dstate = _prep(state), _prep(dstate)
    work = torch.empty(6 * state.numel(), dtype=torch.float64, device=state.device)
    _lib.check(_lib.load().bssn_evolve_tangent(ctypes.byref(_lib.c_params(params)),
                                               ctypes.byref(_lib.c_grid(state.shape[1:])), int(nsteps),
                                               _ptr(state), _ptr(dstate), _ptr(work), stream_handle()))
    return state, dstate


def hamiltonian_l2_tangent(state, dstate, params):
    """(L, dL) with L = sqrt(mean H^2) over all points and dL its derivative
    along dstate (mean(H dH) / L)."""
    shape = tuple(state.shape[1:])
    H = torch.empty(shape, dtype=torch.float64, device=state.device)
    dH = torch.empty_like(H)
    _lib.check(_lib.load().bssn_hamiltonian_tangent(ctypes.byref(_lib.c_params(params)),
                                                    ctypes.byref(_lib.c_grid(shape)), _ptr(state), _ptr(dstate),
                                                    _ptr(H), _ptr(dH), stream_handle()))
    L = torch.sqrt(torch.mean(H * H))
    dL = torch.mean(H * dH) / L
    return float(L), float(dL)


class HamiltonianL2AfterSteps(torch.autograd.Function):
    """L2 Hamiltonian constraint after ``nsteps`` RK4 steps of the initial
    data ``initial(theta)``, differentiable in theta.

    ``initial(theta)`` (theta a NumPy vector) returns the packed initial
    state and, for every component of theta, the packed derivative of the
    initial state with respect to it.  The forward pass runs one tangent
    evolution per component; backward multiplies the stored gradient by the
    incoming gradient.
    """

    @staticmethod
    def forward(ctx, theta, initial, params, nsteps):
        u0, tangents = initial(theta.detach().cpu().numpy())
        value, grads = None, []
        for du0 in tangents:
            u, du = evolve_tangent(u0, du0, params, nsteps)
            value, dL = hamiltonian_l2_tangent(u, du, params)
            grads.append(dL)
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
