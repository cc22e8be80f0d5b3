"""Throughput of the fused BSSN RK4 step on one MI355X.

Workload (the single-puncture configuration): Brill-Lindquist puncture M = 1
at the origin, zero momentum, 256^3 cell-centred grid on [-8, 8]^3, FP64,
1+log slicing, Gamma-driver shift (eta = 2, g = 0.75), kappa = 0, KO
nu = 0.1, Sommerfeld faces, dt = dx / 4.  Initial data is generated on the
device (synthetic: closed-form psi = 1 + M / 2r, W = psi^-2, alpha = W).

A step is one full RK4 step (4 fused stage launches + the Sommerfeld shell
launches).  Reported: cell-updates/s (interior points x steps / time), the
FP32-RHS variant, the HBM roofline fraction of the tile kernel from HIP-event
timings of its launches, and the FP64 NumPy oracle timed on the host.
"""

import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12          # B/s, MI355X HBM3E
FP64_PEAK = 78.6e12        # FLOP/s, MI355X FP64 vector
BYTES_PER_CELL = {1: 3 * 21 * 8, 2: 5 * 21 * 8, 3: 5 * 21 * 8, 4: 3 * 21 * 8}  # algorithmic, per stage

# Algorithmic FP64 operations (add/sub, mul, div, sqrt, cbrt one each; a
# fused multiply-add two), counted by tests/hostcheck/flop_count.cpp:
# the per-point RHS (kappa = 0, vacuum) with the reference's stencil rows as
# dot products over their non-zero taps, at an interior point; one
# algebraic projection; the RK4 combinations of the 21 components (two
# axpys per component in stages 1-3, one in stage 4).  An RK4 step makes
# four RHS evaluations and five projections (stage inputs 2-4, and the
# output projected twice where consecutive steps meet).
FLOPS_RHS = 5344
FLOPS_PROJECTION = 82
FLOPS_STAGE_UPDATES = 21 * (4 + 4 + 4 + 2)
F_ALG = 4 * FLOPS_RHS + 5 * FLOPS_PROJECTION + FLOPS_STAGE_UPDATES  # per cell-update


def puncture_state(n, L=8.0, mass=1.0, device="cuda"):
    import torch
    dx = 2 * L / n
    x = -L + (torch.arange(n, dtype=torch.float64, device=device) + 0.5) * dx
    X, Y, Z = torch.meshgrid(x, x, x, indexing="ij")
    r = torch.sqrt(X * X + Y * Y + Z * Z)
    W = (1.0 + mass / (2.0 * r)) ** -2
    s = torch.zeros((21, n, n, n), dtype=torch.float64, device=device)
    s[0] = 1.0
    s[3] = 1.0
    s[5] = 1.0
    s[6] = W
    s[17] = W
    return s, dx, -L + 0.5 * dx


def puncture_params(P, dx, x0):
    return P(eta=2.0, kappa=0.0, nu=0.1, g=0.75, dx=dx, dt=0.25 * dx, zero_shift=0, gauge=1,
             xl_bc=1, xr_bc=1, yl_bc=1, yr_bc=1, zl_bc=1, zr_bc=1, x_min=x0, y_min=x0, z_min=x0)


def time_steps(J, state, params, steps, warmup, precision):
    import torch
    ev = J.Evolver(state, params, precision=precision)
    ev.step(warmup)
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    ev.step(steps)
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / 1e3, ev


def kappa_step(J, state, params, steps, warmup, ms_kappa0, kappa=0.002):
    """The same puncture step with constraint damping kappa = 0.002 (the
    momentum-constraint term of the A RHS): ms per step and the ratio to the
    kappa = 0 step."""
    import torch
    t, ev = time_steps(J, state.clone(), params._replace(kappa=kappa), steps, warmup, "f64")
    finite = bool(torch.isfinite(ev.state).all())
    del ev
    ms = 1e3 * t / steps
    n3 = state[0].numel()
    return {"kappa": kappa, "ms_per_step": ms, "value": n3 * steps / t, "unit": "cell-updates/s",
            "ratio_vs_kappa0": ms / ms_kappa0, "finite": finite}


def constraints_timing(J, state, params, reps=5):
    """bssn_constraints (H, M_i, det - 1, tr A, Gamma constraint) on the
    workload state: device ms per call and the L2 Hamiltonian constraint."""
    import ctypes

    import torch
    from jax_bssn_amd import _lib
    from jax_bssn_amd.time_evolve import _ptr, stream_handle
    n3 = tuple(state.shape[1:])
    out = torch.empty((9,) + n3, dtype=torch.float64, device=state.device)
    lib = _lib.load()
    cp, cg = _lib.c_params(params), _lib.c_grid(n3)

    def call():
        _lib.check(lib.bssn_constraints(ctypes.byref(cp), ctypes.byref(cg), _ptr(state), _ptr(out), stream_handle()))
    call()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        call()
    e1.record()
    torch.cuda.synchronize()
    return {"ms_per_call": e0.elapsed_time(e1) / reps, "calls": reps,
            "hamiltonian_l2": float(torch.sqrt(torch.mean(out[0] * out[0])))}


# the interior tile kernel, and the face-tile kernels (one per slab with
# closures along its axis, CLA = 1, 2, 4, and the edges and corners, CLA = 7)
KERNEL_NAMES = {("f64", 4): "k_stage_tile<double, 4, 8, 8, false, false, 7>",
                ("f64", 8): "k_stage_tile<double, 4, 8, 8, false, true, CLA> (CLA = 1, 2, 4, 7)",
                ("f32rhs", 4): "k_stage_tile<float, 4, 8, 8, false, false, 7>",
                ("f32rhs", 8): "k_stage_tile<float, 4, 8, 8, false, true, CLA> (CLA = 1, 2, 4, 7)"}
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "k_stage_tile_puncture256_traffic.md")


def measured_traffic(kernel, n):
    """HBM bytes per launch of `kernel` at n^3 from the committed counter
    summary (a line 'traffic_bytes_per_launch <kernel> <n> <bytes>'), else None."""
    try:
        lines = open(TRAFFIC_FILE).read().splitlines()
    except OSError:
        return None
    for ln in lines:
        f = ln.strip("` ").split("\t")
        if len(f) == 4 and f[0] == "traffic_bytes_per_launch" and f[1] == kernel and int(f[2]) == n:
            return float(f[3])
    return None


def kernel_roofline(J, ev, precision, part, n, reps=3):
    """HBM roofline of one tile kernel: the algorithmic bytes of the points it
    updates in each of the four stages over its mean launch duration, timed
    with HIP events recorded on the launch stream."""
    import torch

    from jax_bssn_amd import _lib
    lib = _lib.load()
    cp, cg = ev._cp, ev._cg
    prec = _lib.PREC_F64 if precision == "f64" else _lib.PREC_F32RHS
    npts = int(lib.bssn_stage_points(ctypes.byref(cp), ctypes.byref(cg), part))
    u0 = ev.state
    ustar, ua, ub = ev.work[0], ev.work[1], ev.work[2]
    ua.copy_(u0)
    ustar.copy_(u0)
    total_bytes, total_s, per_stage = 0.0, 0.0, {}
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    for s in (1, 2, 3, 4):
        ms = ctypes.c_float(0.0)
        us = u0 if s == 1 else ua
        _lib.check(lib.bssn_time_stage_part(ctypes.byref(cp), ctypes.byref(cg), prec, s, part,
                                            ctypes.c_void_p(us.data_ptr()), ctypes.c_void_p(u0.data_ptr()),
                                            ctypes.c_void_p(ustar.data_ptr()), ctypes.c_void_p(ub.data_ptr()),
                                            reps, ctypes.byref(ms), ctypes.c_void_p(ev.stage_scratch.data_ptr()),
                                            st))
        per_stage[s] = ms.value
        total_bytes += BYTES_PER_CELL[s] * npts
        total_s += ms.value / 1e3
    achieved = total_bytes / total_s
    kernel = KERNEL_NAMES[(precision, part)]
    traffic = measured_traffic(kernel, n)
    return {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
            "frac": achieved / HBM_PEAK, "traffic": traffic, "kernel": kernel,
            "launch_ms_by_stage": per_stage, "launch_ms_avg": 1e3 * total_s / 4,
            "algorithmic_bytes_per_launch_avg": total_bytes / 4, "points_per_launch": npts}


def f32_accuracy(s64, s32, params, steps):
    """Error of the FP32-RHS state against the FP64 state after the same
    steps from the same data: per-field max-norm relative difference and the
    L2 Hamiltonian constraint of both."""
    from jax_bssn_amd.constraints import NORM_NAMES, constraints_packed, norms_packed
    groups = {"conformal_metric": (0, 6), "conformal_factor": (6, 7), "traceless_K": (7, 13), "trace_K": (13, 14),
              "conformal_connection": (14, 17), "lapse": (17, 18), "shift": (18, 21)}
    rel = {}
    for name, (a, b) in groups.items():
        scale = float(s64[a:b].abs().max())
        rel[name] = float((s32[a:b] - s64[a:b]).abs().max()) / max(scale, 1e-300)
    h = [dict(zip(NORM_NAMES, norms_packed(constraints_packed(s, params)).cpu().tolist()))["hamiltonian_l2"]
         for s in (s64, s32)]
    return {"steps_compared": steps, "max_rel_diff_vs_f64_by_field": rel, "hamiltonian_l2_f64": h[0],
            "hamiltonian_l2_f32rhs": h[1], "hamiltonian_l2_ratio": h[1] / h[0]}


def batched_gauge_waves(J, steps, count=64, seed=0, precision="f64"):
    """64 periodic gauge waves of edge n in {32, 48, 64, 80, 96}, A ~ U(0.01,
    0.1), d ~ U(0.5, 1), dx = d / n, KO nu = 0.1, dt = dx / 4: one batched,
    autotuned, graph-captured RK4 step per replay.  Returns aggregate
    cell-updates/s, the rate of the same instances evolved one grid at a
    time and the largest per-instance difference to an instance evolved
    alone."""
    import numpy as np
    import torch

    from jax_bssn_amd.batch import BatchEvolver
    from jax_bssn_amd.initial_data import gauge_wave_analytic_state, periodic_coordinate_arrays
    from jax_bssn_amd.variables import pack
    rng = np.random.default_rng(seed)
    states, params = [], []
    for _ in range(count):
        n = int(rng.choice([32, 48, 64, 80, 96]))
        amp, d = rng.uniform(0.01, 0.1), rng.uniform(0.5, 1.0)
        X, Y, Z = periodic_coordinate_arrays(n, d)
        states.append(pack(gauge_wave_analytic_state(X, Y, Z, amplitude=amp, wavelength=d)))
        params.append(J.BSSNParameters(eta=0.0, kappa=0.0, g=0.0, nu=0.1, dx=d / n, dt=0.25 * d / n,
                                       zero_shift=1, gauge=0))
    cells = sum(int(s[0].numel()) for s in states)
    b = BatchEvolver([s.clone() for s in states], params, precision=precision)
    b.capture()  # one warm-up step
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    b.step(steps)
    e1.record()
    torch.cuda.synchronize()
    t = e0.elapsed_time(e1) / 1e3
    worst = 0.0
    for i in (0, count // 2, count - 1):  # spot-check instances against unbatched runs
        ev = J.Evolver(states[i].clone(), params[i], precision=precision)
        ev.step(steps + 1)
        worst = max(worst, float((b.states[i] - ev.state).abs().max() / ev.state.abs().max()))
    # the same instances evolved one grid at a time (one warm-up step each)
    evs = [J.Evolver(s.clone(), p, precision=precision).step(1) for s, p in zip(states, params)]
    torch.cuda.synchronize()
    e0.record()
    for ev in evs:
        ev.step(steps)
    e1.record()
    torch.cuda.synchronize()
    t1 = e0.elapsed_time(e1) / 1e3
    del evs
    return {"value": cells * steps / t, "unit": "cell-updates/s", "instances": count, "cells": cells,
            "ms_per_step": 1e3 * t / steps, "precision": precision, "tuning": b.tuning,
            "unbatched_value": cells * steps / t1, "speedup_vs_unbatched": t1 / t,
            "max_rel_diff_vs_unbatched": worst}


def batched_punctures(J, steps, count=16, seed=1):
    """16 single punctures (Sommerfeld faces, 1+log, Gamma driver, the
    workload's parameters) of edge n in {32, 48, 64, 80, 96} on [-8, 8]^3
    with masses M ~ U(0.5, 1.5): one batched, autotuned, graph-captured RK4
    step per replay (lean and face tiles in their own launches).
    Returns aggregate cell-updates/s and the largest per-instance difference
    to the same instance evolved alone."""
    import numpy as np
    import torch

    from jax_bssn_amd.batch import BatchEvolver
    rng = np.random.default_rng(seed)
    states, params = [], []
    for _ in range(count):
        n = int(rng.choice([32, 48, 64, 80, 96]))
        s, dx, x0 = puncture_state(n, mass=float(rng.uniform(0.5, 1.5)))
        states.append(s)
        params.append(puncture_params(J.BSSNParameters, dx, x0))
    cells = sum(int(s[0].numel()) for s in states)
    b = BatchEvolver([s.clone() for s in states], params)
    b.capture()  # one warm-up step
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    b.step(steps)
    e1.record()
    torch.cuda.synchronize()
    t = e0.elapsed_time(e1) / 1e3
    worst = 0.0
    for i in (0, count // 2, count - 1):
        ev = J.Evolver(states[i].clone(), params[i])
        ev.step(steps + 1)
        worst = max(worst, float((b.states[i] - ev.state).abs().max() / ev.state.abs().max()))
    # the same instances evolved one grid at a time (one warm-up step each)
    evs = [J.Evolver(s.clone(), p).step(1) for s, p in zip(states, params)]
    torch.cuda.synchronize()
    e0.record()
    for ev in evs:
        ev.step(steps)
    e1.record()
    torch.cuda.synchronize()
    t1 = e0.elapsed_time(e1) / 1e3
    del evs
    return {"value": cells * steps / t, "unit": "cell-updates/s", "instances": count, "cells": cells,
            "ms_per_step": 1e3 * t / steps, "precision": "f64", "tuning": b.tuning,
            "face_tiles": int(b.plan.face_tiles), "lean_tiles": int(b.plan.total_tiles),
            "unbatched_value": cells * steps / t1, "speedup_vs_unbatched": t1 / t,
            "max_rel_diff_vs_unbatched": worst}


def fmr_linear_wave(J, steps, n=64, levels=3, amplitude=1.0e-8):
    """Linearized plane wave (h+ = 1e-8 sin 2 pi x, wavelength 1) through a
    nested FMR hierarchy as in the reference's linear-wave demo
    (demos/gravitational_waves/linear_wave.py: root n^3 periodic centred at
    the origin, each child the centred half of its parent's active shape,
    MAD weights, all levels stepped with the finest dt): aggregate
    cell-updates/s over the active points of every level."""
    import numpy as np
    import torch

    from jax_bssn_amd import fmr as F
    from jax_bssn_amd.initial_data import linear_wave_initial_data
    from jax_bssn_amd.variables import pack
    dx = 1.0 / n
    dt = 0.25 * dx / 2 ** (levels - 1)
    origin = -(n - 1) * dx / 2.0
    base = J.BSSNParameters(eta=0.0, kappa=0.0, nu=0.0, g=0.0, dx=dx, dt=dt, zero_shift=1, gauge=0,
                            x_min=origin, y_min=origin, z_min=origin)
    patches, active, params = [], [(n,) * 3], [base]
    for lv in range(1, levels):
        parent = params[-1]
        pg = 0 if lv == 1 else 4  # guard points before the parent's active vertex 0
        lo = tuple((s - s // 2) // 2 for s in active[-1])
        hi = tuple(l + s // 2 - 1 for l, s in zip(lo, active[-1]))
        patches.append(F.FMRPatchSpec(lo, hi))
        active.append(tuple(2 * (h - l) + 1 for l, h in zip(lo, hi)))
        cdx = parent.dx / 2
        params.append(parent._replace(dx=cdx, x_min=parent.x_min + (pg + lo[0]) * parent.dx - 4 * cdx,
                                      y_min=parent.y_min + (pg + lo[1]) * parent.dx - 4 * cdx,
                                      z_min=parent.z_min + (pg + lo[2]) * parent.dx - 4 * cdx))
    states = []
    for lv, p in enumerate(params):
        shape = active[lv] if lv == 0 else tuple(a + 8 for a in active[lv])
        x = p.x_min + p.dx * np.arange(shape[0])
        X = np.broadcast_to(x[:, None, None], shape)
        states.append(pack(linear_wave_initial_data(X, X, X, amplitude=amplitude)))
    ev = F.FMREvolver(states, params, F.FMRHierarchySpec(tuple(patches), use_mad=True))
    ev.step(1)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    ev.step(steps)
    e1.record()
    torch.cuda.synchronize()
    t = e0.elapsed_time(e1) / 1e3
    cells = sum(int(np.prod(a)) for a in active)
    # h+ = (gamma~_yy - gamma~_zz) / (2 W^2) on the active points of every
    # level against the plane wave A sin 2 pi (x - t), relative to A
    time_now = (steps + 1) * dt
    errors = []
    for lv, (u, p) in enumerate(zip(ev.u, params)):
        g = 0 if lv == 0 else 4
        a = u[:, g:u.shape[1] - g, g:u.shape[2] - g, g:u.shape[3] - g]
        hp = 0.5 * (a[3] - a[5]) / a[6] ** 2
        x = p.x_min + p.dx * (g + torch.arange(a.shape[1], dtype=torch.float64, device=a.device))
        exact = amplitude * torch.sin(2 * np.pi * (x - time_now))[:, None, None]
        errors.append(float((hp - exact).abs().max()) / amplitude)
    return {"value": cells * steps / t, "unit": "cell-updates/s (active points, all levels)",
            "levels": levels, "active_shapes": [list(a) for a in active], "ms_per_step": 1e3 * t / steps,
            "steps": steps + 1, "finite": bool(all(torch.isfinite(u).all() for u in ev.u)),
            "h_plus_max_error_over_amplitude_by_level": errors}


def axisymmetric_puncture(J, steps, nr=512, nz=1024, separation=1.5, precision="f64", keep_state=False):
    """Head-on pair of zero-momentum punctures (m = 0.5 at z = +-s/2) on the
    compact (rho, z) plane nr x nz (+ 4 parity ghosts), dx = 1/32, moving-
    puncture gauge (1+log, Gamma driver eta = 2, kappa = 0.002, KO nu = 0.02),
    Sommerfeld outer faces: cell-updates/s of compact points."""
    import numpy as np
    import torch

    from jax_bssn_amd.cartoon import AxisymmetricEvolver
    from jax_bssn_amd.initial_data import puncture_state
    dx = 1.0 / 32
    rho = (np.arange(-4, nr) + 0.5) * dx
    z = (np.arange(nz) - (nz - 1) / 2.0) * dx
    R, Z = np.meshgrid(rho, z, indexing="ij")
    X, Y = R[:, None, :], np.zeros_like(R[:, None, :])
    v = puncture_state(X, Y, Z[:, None, :], masses=(0.5, 0.5),
                       positions=((0.0, 0.0, -separation / 2), (0.0, 0.0, separation / 2)))
    p = J.BSSNParameters(eta=2.0, kappa=0.002, nu=0.02, g=0.75, dx=dx, dt=0.2 * dx, zero_shift=0, gauge=1,
                         xl_bc=0, xr_bc=1, yl_bc=0, yr_bc=0, zl_bc=1, zr_bc=1, x_min=-3.5 * dx, y_min=-4.0 * dx,
                         z_min=float(z[0]))
    ev = AxisymmetricEvolver(v, p, precision=precision)
    ev.step(1)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    ev.step(steps)
    e1.record()
    torch.cuda.synchronize()
    t = e0.elapsed_time(e1) / 1e3
    out = {"value": nr * nz * steps / t, "unit": "cell-updates/s (compact points)", "compact": [nr, nz],
           "ms_per_step": 1e3 * t / steps, "precision": precision, "finite": bool(torch.isfinite(ev.u).all())}
    if keep_state:
        out["state"] = ev.u
    return out


def tangent_gradient(J, n=128, steps=16):
    """Gradient of the L2 Hamiltonian constraint after ``steps`` RK4 steps of
    the single puncture (n^3 on [-8, 8]^3, as the main workload) with respect
    to the mass M and a Gaussian lapse amplitude eps (width 1, centre drawn
    from default_rng(0)), by one forward-mode tangent run carrying both
    directions.  The device time of that run plus the Hamiltonian tangent is
    set against the plain forward evolution of the same steps (objective
    excluded); the end-to-end autograd call (initial data on the host
    included) is reported beside it."""
    import numpy as np
    import torch

    from jax_bssn_amd.tangent import (HamiltonianL2AfterSteps, evolve_tangent, hamiltonian_l2_tangent,
                                      puncture_initial)
    from jax_bssn_amd.time_evolve import evolve_packed
    L = 8.0
    dx = 2 * L / n
    x = -L + (np.arange(n) + 0.5) * dx
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    c = np.random.default_rng(0).uniform(-1, 1, 3)
    initial = puncture_initial(X, Y, Z, width=1.0, lapse_centre=tuple(c))
    p = puncture_params(J.BSSNParameters, dx, x[0])
    theta0 = np.array([1.0, 1e-3])
    u0, tangents = initial(theta0)
    du0 = torch.stack(tangents)

    def timed(fn):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        out = fn()
        torch.cuda.synchronize()
        return time.perf_counter() - t0, out

    timed(lambda: evolve_packed(u0.clone(), p, 1))
    timed(lambda: evolve_tangent(u0, du0, p, 1))
    t_fwd, _ = timed(lambda: evolve_packed(u0.clone(), p, steps))

    def value_and_gradient():
        u, du = evolve_tangent(u0, du0, p, steps)
        return hamiltonian_l2_tangent(u, du, p)
    t_grad, (val, grad) = timed(value_and_gradient)
    theta = torch.tensor(theta0, requires_grad=True)
    t_e2e, obj = timed(lambda: HamiltonianL2AfterSteps.apply(theta, initial, p, steps))
    obj.backward()

    # central differences of the forward objective at the same size
    from jax_bssn_amd.constraints import constraints_packed

    def objective(th):
        u, _ = initial(th)
        evolve_packed(u, p, steps)
        h = constraints_packed(u, p)[0]
        return float(torch.sqrt(torch.mean(h * h)))
    fd = []
    for k, h in ((0, 1e-4), (1, 1e-4)):
        tp, tm = theta0.copy(), theta0.copy()
        tp[k] += h
        tm[k] -= h
        fd.append((objective(tp) - objective(tm)) / (2 * h))
    return {"method": "forward-mode tangent (dual numbers), both parameters in one run", "grid": [n, n, n],
            "steps": steps, "parameters": ["M", "eps"], "hamiltonian_l2": val, "gradient": list(grad),
            "autograd_gradient": theta.grad.tolist(), "central_difference_gradient": fd,
            "rel_diff_vs_central_difference": [abs(g - f) / abs(f) for g, f in zip(grad, fd)],
            "forward_seconds": t_fwd,
            "value_and_gradient_seconds": t_grad, "gradient_forward_ratio": t_grad / t_fwd,
            "autograd_call_seconds": t_e2e}


def adjoint_gradient(J, n=128, steps=16, checkpoint_every=4, forward_mode=None):
    """The same gradient as ``tangent_gradient`` (L2 Hamiltonian constraint
    after ``steps`` RK4 steps of the single puncture with respect to M and
    eps) by the reverse sweep: the forward evolution storing a checkpoint
    every ``checkpoint_every`` steps, the adjoint of the objective, then the
    steps backwards (each segment recomputed from its checkpoint), the
    gradient being the initial adjoint against the initial-data
    derivatives.  Timed against the plain forward evolution of the same
    steps and compared with the forward-mode gradient."""
    import numpy as np
    import torch

    from jax_bssn_amd.adjoint import _Ops, evolve_vjp, hamiltonian_l2_vjp
    from jax_bssn_amd.tangent import puncture_initial
    from jax_bssn_amd.time_evolve import evolve_packed

    L = 8.0
    dx = 2 * L / n
    x = -L + (np.arange(n) + 0.5) * dx
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    c = np.random.default_rng(0).uniform(-1, 1, 3)
    initial = puncture_initial(X, Y, Z, width=1.0, lapse_centre=tuple(c))
    p = puncture_params(J.BSSNParameters, dx, x[0])
    u0, tangents = initial(np.array([1.0, 1e-3]))

    def timed(fn):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        out = fn()
        torch.cuda.synchronize()
        return time.perf_counter() - t0, out

    ops = _Ops(u0, p)
    timed(lambda: evolve_packed(u0.clone(), p, 1))
    t_fwd, _ = timed(lambda: evolve_packed(u0.clone(), p, steps))

    def value_and_gradient():
        ubar, value = evolve_vjp(u0, steps, p, lambda u: hamiltonian_l2_vjp(u, p, ops), checkpoint_every, ops)
        return value, [float(torch.sum(ubar * t)) for t in tangents]
    t_grad, (val, grad) = timed(value_and_gradient)
    out = {"method": f"reverse-mode adjoint, checkpoint every {checkpoint_every} steps, segments recomputed",
           "grid": [n, n, n], "steps": steps, "parameters": ["M", "eps"], "hamiltonian_l2": val,
           "gradient": grad, "forward_seconds": t_fwd, "value_and_gradient_seconds": t_grad,
           "gradient_forward_ratio": t_grad / t_fwd}
    if forward_mode:
        fm = forward_mode["gradient"]
        out["rel_diff_vs_forward_mode"] = [abs(a - b) / abs(b) for a, b in zip(grad, fm)]
    return out


def copy_calibration(nbytes=1 << 31, reps=3):
    """A plain device copy of known size (read nbytes, write nbytes), for
    scaling the FETCH_SIZE / WRITE_SIZE counters to bytes."""
    import torch
    src = torch.ones(nbytes // 8, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(reps):
        torch.neg(src, out=dst)
    torch.cuda.synchronize()
    del src, dst


def _cpu_model():
    try:
        for ln in open("/proc/cpuinfo"):
            if ln.startswith("model name"):
                return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    import platform
    return platform.processor() or "unknown"


def cpu_baseline(sizes=(32, 48, 64), steps=3):
    """FP64 NumPy oracle on the host, one thread: the same puncture
    configuration at each size, one warm-up step and then the median of
    ``steps`` timed RK4 steps.  ``value`` is the rate at the largest size."""
    import numpy as np
    from threadpoolctl import threadpool_limits

    from oracle import bssn as OB
    from oracle import initial_data as OI
    L = 8.0
    per_size = {}
    with threadpool_limits(limits=1):
        for n in sizes:
            dx = 2 * L / n
            x = -L + (np.arange(n) + 0.5) * dx
            X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
            v = OI.puncture_state(X, Y, Z, masses=(1.0,), positions=((0.0, 0.0, 0.0),))
            p = puncture_params(OB.BSSNParameters, dx, x[0])
            v = OB.rk4_step(v, p)  # warm-up
            times = []
            for _ in range(steps):
                t0 = time.perf_counter()
                v = OB.rk4_step(v, p)
                times.append(time.perf_counter() - t0)
            med = float(np.median(times))
            per_size[str(n)] = {"median_step_seconds": med, "cell_updates_per_s": n ** 3 / med}
            print(f"[cpu baseline] {n}^3: median step {med:.2f} s", file=sys.stderr, flush=True)
    n = max(sizes)
    threads = {k: os.environ.get(k) for k in ("OMP_NUM_THREADS", "OPENBLAS_NUM_THREADS", "MKL_NUM_THREADS")}
    return {"value": per_size[str(n)]["cell_updates_per_s"], "unit": "cell-updates/s", "cores": 1,
            "kind": "port",
            "sample": f"median of {steps} RK4 steps (after 1 warm-up) of the same puncture configuration at "
                      f"{', '.join(f'{m}^3' for m in sizes)} on [-8, 8]^3; value at {n}^3; NumPy FP64 oracle "
                      "limited to 1 thread (threadpoolctl)",
            "per_size": per_size, "cpu_model": _cpu_model(), "host_cores": len(os.sched_getaffinity(0)),
            "numpy": np.__version__, "thread_env": threads}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--size", type=int, default=256)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-f32", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-cartoon", action="store_true")
    ap.add_argument("--no-gradient", action="store_true")
    ap.add_argument("--no-fmr", action="store_true")
    ap.add_argument("--no-kappa", action="store_true",
                    help="skip the kappa = 0.002 step and the constraint-kernel legs")
    ap.add_argument("--accuracy-steps", type=int, default=50,
                    help="steps of the FP64 / FP32-RHS comparison run (0: none)")
    ap.add_argument("--adjoint-size", type=int, default=0,
                    help="grid edge of the reverse-mode gradient leg (0: --size)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run a 2 GiB device copy (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); the multi-GPU "
              "halo-exchange path is design-only in this build", file=sys.stderr)
        sys.exit(2)

    import torch

    import jax_bssn_amd as J
    torch.cuda.set_device(int(os.environ.get("LOCAL_RANK", 0)))
    n = args.size
    state, dx, x0 = puncture_state(n)
    params = puncture_params(J.BSSNParameters, dx, x0)
    cells = n ** 3

    t64, ev64 = time_steps(J, state.clone(), params, args.steps, args.warmup, "f64")
    rate64 = cells * args.steps / t64
    final64 = ev64.state.clone()
    roof = kernel_roofline(J, ev64, "f64", 4, n)
    closure = kernel_roofline(J, ev64, "f64", 8, n)
    del ev64
    extra = {}
    if not args.no_f32:
        t32, ev32 = time_steps(J, state.clone(), params, args.steps, args.warmup, "f32rhs")
        rate32 = cells * args.steps / t32
        acc = f32_accuracy(final64, ev32.state, params, args.warmup + args.steps)
        roof32 = kernel_roofline(J, ev32, "f32rhs", 4, n)
        del ev32
        extra["f32rhs"] = {"value": rate32, "ms_per_step": 1e3 * t32 / args.steps,
                           "speedup_vs_f64": rate32 / rate64, "tile_kernel_hbm_frac": roof32["frac"], **acc}
        if args.accuracy_steps > 0:
            # the full 50-step run of the puncture configuration in both precisions
            e64 = J.Evolver(state.clone(), params, precision="f64").step(args.accuracy_steps)
            e32 = J.Evolver(state.clone(), params, precision="f32rhs").step(args.accuracy_steps)
            extra["f32rhs"]["long_run"] = f32_accuracy(e64.state, e32.state, params, args.accuracy_steps)
            del e64, e32
    if not args.no_kappa:
        extra["kappa_step"] = kappa_step(J, state, params, args.steps, args.warmup, 1e3 * t64 / args.steps)
        extra["constraints"] = constraints_timing(J, final64, params)
    del final64
    torch.cuda.synchronize()
    line = {
        "metric": "BSSN RK4 cell-updates/s at 256^3 FP64 (single puncture)",
        "value": rate64,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * t64 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (closed-form Brill-Lindquist puncture generated on the device)",
        "config": {"workload": f"single puncture M=1, {n}^3 on [-8,8]^3, 1+log + Gamma driver, "
                               "Sommerfeld faces, KO nu=0.1, dt=dx/4, FP64 RK4",
                   "grid": [n, n, n], "global_batch": 1, "parallelism": "single GPU"},
        "roofline": roof,
        # the whole step against the FP64 vector peak
        "fp64_compute": {"flops_per_cell_update": F_ALG, "achieved": rate64 * F_ALG / 1e12,
                         "peak": FP64_PEAK / 1e12, "unit": "TFLOP/s", "frac": rate64 * F_ALG / FP64_PEAK},
        "closure_tile_kernel": {k: closure[k] for k in ("kernel", "achieved", "frac", "launch_ms_avg",
                                                         "points_per_launch")},
        # gradient of the L2 Hamiltonian constraint w.r.t. two initial-data
        # parameters over its time of the forward evolution (filled below)
        "gradient_forward_ratio": None,
        "gradient_method": None,
    }
    if not args.no_batch:
        extra["batched_gauge_waves"] = batched_gauge_waves(J, args.steps)
        extra["batched_gauge_waves_f32rhs"] = batched_gauge_waves(J, args.steps, precision="f32rhs")
        extra["batched_punctures"] = batched_punctures(J, args.steps)
    if not args.no_cartoon:
        a64 = axisymmetric_puncture(J, args.steps, keep_state=True)
        a32 = axisymmetric_puncture(J, args.steps, precision="f32rhs", keep_state=True)
        s64, s32 = a64.pop("state"), a32.pop("state")
        a32["max_rel_diff_vs_f64"] = float((s32 - s64).abs().max() / s64.abs().max())
        a32["speedup_vs_f64"] = a32["value"] / a64["value"]
        del s64, s32
        extra["axisymmetric_puncture"] = a64
        extra["axisymmetric_puncture_f32rhs"] = a32
    if not args.no_fmr:
        extra["fmr_linear_wave"] = fmr_linear_wave(J, args.steps)
    if not args.no_gradient:
        extra["tangent_gradient"] = tangent_gradient(J, n=n)
        extra["adjoint_gradient"] = adjoint_gradient(J, n=args.adjoint_size or n, forward_mode=(
            extra["tangent_gradient"] if (args.adjoint_size or n) == n else None))
        line["gradient_forward_ratio"] = extra["adjoint_gradient"]["gradient_forward_ratio"]
        line["gradient_method"] = extra["adjoint_gradient"]["method"]
    line.update(extra)
    if args.copy_calibration:
        copy_calibration()
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
