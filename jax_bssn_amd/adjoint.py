"""Reverse-mode differentiation (the adjoint) of the RK4 evolution, with
checkpoints, and a torch.autograd.Function for gradients of the L2
Hamiltonian constraint.

The reference differentiates its JAX evolution with jax.grad.  Here the
reverse sweep of ``rk4_step`` (reference evolution/time_evolve.py:144-217)
is composed from the library's transposed linearisations (bssn_rhs_vjp:
J_F^T of the RHS and J_H^T of the Hamiltonian constraint; bssn_project_vjp:
J_P^T of the algebraic projection), kappa = 0, vacuum, whole grids.  One
step, with u0' = P(u):

    k1 = F(u0'),  m_s = P(u0' + c_s dt k_s),  k_{s+1} = F(m_s)  (c = 1/2, 1/2, 1)
    u_new = P(u0' + dt/6 (k1 + 2 k2 + 2 k3 + k4))

and its adjoint runs the stages backwards from the adjoint of u_new,
recomputing the step's stage states from u first.  ``evolve_vjp`` stores
the state every ``checkpoint_every`` steps on the way forward and, segment
by segment backwards, recomputes the states of a segment from its
checkpoint before sweeping back through it: memory of one segment plus one
step's stages, about one extra forward evolution of work.
"""

import ctypes

import torch

from . import _lib
from .time_evolve import _ptr, _scratch, evolve_packed, require_packed, stream_handle


def _params_grid(state, params):
    return _lib.c_params(params), _lib.c_grid(tuple(state.shape[1:]))


class _Ops:
    """The forward and transposed operators of one grid shape and parameter
    set on device buffers (work buffers allocated once)."""

    def __init__(self, state, params):
        require_packed(state)
        self.params = params
        self.cp, self.cg = _params_grid(state, params)
        lib = _lib.load()
        self.lib = lib
        self.work_rhs = torch.empty(max(_lib.work_size(params, tuple(state.shape[1:]), 0), 1), dtype=torch.float64,
                                    device=state.device)
        size = int(lib.bssn_vjp_work_size(ctypes.byref(self.cg)))
        if size < 0:
            _lib.check(size)
        self.work_vjp = torch.empty(size, dtype=torch.float64, device=state.device)

    def F(self, u):
        out = torch.empty_like(u)
        _lib.check(self.lib.bssn_rhs(ctypes.byref(self.cp), ctypes.byref(self.cg), _lib.PREC_F64, _ptr(u), _ptr(out),
                                     _ptr(self.work_rhs), stream_handle()))
        return out

    def P(self, u):
        out = torch.empty_like(u)
        _lib.check(self.lib.bssn_enforce_constraints(ctypes.byref(self.cg), _ptr(u), _ptr(out), stream_handle()))
        return out

    def FT(self, u, kbar, out=None, accumulate=False):
        """out (=, or += with accumulate) J_F(u)^T kbar."""
        if out is None:
            out, accumulate = torch.empty_like(u), False
        _lib.check(self.lib.bssn_rhs_vjp(ctypes.byref(self.cp), ctypes.byref(self.cg), _ptr(u), _ptr(kbar), None,
                                         _ptr(out), int(bool(accumulate)), _ptr(self.work_vjp), stream_handle()))
        return out

    def HT(self, u, hbar):
        out = torch.empty_like(u)
        _lib.check(self.lib.bssn_rhs_vjp(ctypes.byref(self.cp), ctypes.byref(self.cg), _ptr(u), None, _ptr(hbar),
                                         _ptr(out), 0, _ptr(self.work_vjp), stream_handle()))
        return out

    def PT(self, u_pre, ubar):
        """J_P(u_pre)^T ubar, in place."""
        _lib.check(self.lib.bssn_project_vjp(ctypes.byref(self.cg), _ptr(u_pre), _ptr(ubar), stream_handle()))
        return ubar


def rk4_step_vjp(state, state_bar, params, ops=None):
    """Adjoint of one ``rk4_step``: J_step(state)^T state_bar for packed
    device states (state: the input of the step)."""
    ops = ops or _Ops(state, params)
    dt = float(params.dt)
    # (each line one fused elementwise kernel over the state)
    u0 = ops.P(state)
    k1 = ops.F(u0)
    pre1 = torch.add(u0, k1, alpha=0.5 * dt)
    m1 = ops.P(pre1)
    k2 = ops.F(m1)
    pre2 = torch.add(u0, k2, alpha=0.5 * dt)
    m2 = ops.P(pre2)
    k3 = ops.F(m2)
    pre3 = torch.add(u0, k3, alpha=dt)
    m3 = ops.P(pre3)
    k4 = ops.F(m3)
    acc = torch.add(k1, k2, alpha=2.0)
    del k1, k2
    acc.add_(k3, alpha=2.0).add_(k4)
    del k3, k4
    pre_out = torch.add(u0, acc, alpha=dt / 6.0)
    del acc
    pbar = ops.PT(pre_out, state_bar.clone())
    del pre_out
    u0_bar = pbar.clone()
    k1_bar, k4_bar = torch.mul(pbar, dt / 6.0), torch.mul(pbar, dt / 6.0)
    k2_bar, k3_bar = torch.mul(pbar, dt / 3.0), torch.mul(pbar, dt / 3.0)
    # stage 4 <- m3 <- pre3 = u0 + dt k3
    b = ops.PT(pre3, ops.FT(m3, k4_bar))
    u0_bar += b
    k3_bar.add_(b, alpha=dt)
    # stage 3 <- m2 <- pre2 = u0 + dt/2 k2
    b = ops.PT(pre2, ops.FT(m2, k3_bar))
    u0_bar += b
    k2_bar.add_(b, alpha=0.5 * dt)
    # stage 2 <- m1 <- pre1 = u0 + dt/2 k1
    b = ops.PT(pre1, ops.FT(m1, k2_bar))
    u0_bar += b
    k1_bar.add_(b, alpha=0.5 * dt)
    # stage 1 <- u0
    ops.FT(u0, k1_bar, out=u0_bar, accumulate=True)
    return ops.PT(state, u0_bar)


def evolve_vjp(state, nsteps, params, final_bar, checkpoint_every=4, ops=None):
    """Adjoint of ``nsteps`` RK4 steps from ``state``: returns the adjoint
    of the initial state given ``final_bar`` as a function of the final
    state -- or, when ``final_bar`` is callable, of ``final_bar(final
    state)``, which returns (value, adjoint of the final state).  Returns
    (initial adjoint, value or None)."""
    require_packed(state)
    ops = ops or _Ops(state, params)
    work = _scratch(params, tuple(state.shape[1:]), 1, state.device)
    k = max(1, int(checkpoint_every))
    checkpoints = [state.clone()]
    u = state.clone()
    for s in range(nsteps):
        evolve_packed(u, params, 1, work)
        if (s + 1) % k == 0 and s + 1 < nsteps:
            checkpoints.append(u.clone())
    value = None
    if callable(final_bar):
        value, ubar = final_bar(u)
    else:
        ubar = final_bar.clone()
    del u
    for seg in reversed(range(len(checkpoints))):
        first = seg * k
        count = min(k, nsteps - first)
        states = [checkpoints[seg]]
        for _ in range(count - 1):
            nxt = states[-1].clone()
            evolve_packed(nxt, params, 1, work)
            states.append(nxt)
        for s in reversed(range(count)):
            ubar = rk4_step_vjp(states[s], ubar, params, ops)
        del states
        checkpoints.pop()
    return ubar, value


def hamiltonian_l2_vjp(state, params, ops=None):
    """(L, dL/dstate) with L = sqrt(mean H^2) over all points."""
    require_packed(state)
    ops = ops or _Ops(state, params)
    shape = tuple(state.shape[1:])
    out = torch.empty((9,) + shape, dtype=torch.float64, device=state.device)
    _lib.check(ops.lib.bssn_constraints(ctypes.byref(ops.cp), ctypes.byref(ops.cg), _ptr(state), _ptr(out),
                                        stream_handle()))
    H = out[0].contiguous()
    L = torch.sqrt(torch.mean(H * H))
    hbar = (H / (H.numel() * L)).contiguous()
    return float(L), ops.HT(state, hbar)


class HamiltonianL2Reverse(torch.autograd.Function):
    """L2 Hamiltonian constraint after ``nsteps`` RK4 steps of the initial
    data ``initial(theta)``, differentiable in theta by the reverse sweep.

    ``initial(theta)`` (theta a NumPy vector) returns the packed initial
    state and, for every component of theta, the packed derivative of the
    initial state with respect to it (as jax_bssn_amd.tangent's
    puncture_initial).  Forward runs the evolution storing checkpoints and
    the reverse sweep at once; the gradient with respect to every component
    of theta is the inner product of the initial adjoint with that
    component's initial-state derivative, however many components there are.
    """

    @staticmethod
    def forward(ctx, theta, initial, params, nsteps, checkpoint_every=4):
        u0, tangents = initial(theta.detach().cpu().numpy())
        ops = _Ops(u0, params)
        u0_bar, value = evolve_vjp(u0, nsteps, params, lambda u: hamiltonian_l2_vjp(u, params, ops),
                                   checkpoint_every, ops)
        grads = [float(torch.sum(u0_bar * t)) for t in tangents]
        ctx.save_for_backward(torch.tensor(grads, dtype=theta.dtype))
        return torch.tensor(value, dtype=theta.dtype)

    @staticmethod
    def backward(ctx, grad_out):
        (grads,) = ctx.saved_tensors
        return grad_out * grads, None, None, None, None
