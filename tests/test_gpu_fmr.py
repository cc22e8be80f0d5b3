"""GPU parity of the FMR companion kernels and the synchronized RK4 step.

* prolongation / restriction kernels against the oracle's dense matrices
  and injection (oracle/fmr.py) on random data: 1e-13 relative (different
  summation order only), for guard widths 1, 2, 4 and 7 (row, column and
  point kernels);
* one and two hierarchy steps against the oracle restatement of the frozen
  serial algorithm: per-field max-norm relative difference <= 1e-12 per
  level;
* the reference's published known answers (docs/fmr.rst:191-217,
  tests/FMR/test_tiled_fmr_convergence.py): the composite lapse error of the
  interface-crossing gauge wave on 16 x 8 x 8 and 24 x 12 x 12 roots after
  256 steps, and the temporal differences of the 24 x 8 x 8 root at 20, 40
  and 80 steps.
"""

import ctypes
import math

import numpy as np
import pytest
import torch

from oracle import bssn as OB
from oracle import fmr as OF
from oracle import initial_data as OI
from tests.parity_utils import rel_errors, to_np

pytestmark = pytest.mark.gpu

STEP_TOL = 1e-12


@pytest.fixture(scope="module")
def J():
    import jax_bssn_amd as J
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    return J


def _grid(shape):
    from jax_bssn_amd import _lib
    return ctypes.byref(_lib.c_grid(shape))


def _i3(v):
    return (ctypes.c_int32 * 3)(*v)


def _p(t):
    return ctypes.c_void_p(t.data_ptr())


@pytest.mark.parametrize("offset,periodic", [(0, True), (4, False)])
def test_prolongation_kernel_matches_oracle(J, offset, periodic):
    from jax_bssn_amd import _lib
    lib = _lib.load()
    rng = np.random.default_rng(11 + offset)
    spec = OF.FMRPatchSpec((3, 4, 3), (7, 7, 8)) if periodic else OF.FMRPatchSpec((2, 3, 2), (6, 6, 7))
    pshape = (12, 11, 13) if periodic else (12 + 2 * offset, 11 + 2 * offset, 13 + 2 * offset)
    cshape = tuple(n + 8 for n in OF.fine_active_shape(spec))
    parent = rng.standard_normal((21,) + pshape)
    child = rng.standard_normal((21,) + cshape)
    mats = OF.interpolation_matrices(pshape, spec, offset, periodic)
    act = (slice(None),) + OF.fine_active_slice(spec)
    want = OF.prolongate(parent, mats)
    want[act] = child[act]
    gp = torch.tensor(parent, device="cuda")
    gc = torch.tensor(child, device="cuda")
    _lib.check(lib.bssn_fmr_prolong(_grid(pshape), offset, _i3((int(periodic),) * 3), _p(gp), _grid(cshape),
                                    _i3(spec.coarse_lo), 4, _p(gc), None))
    got = gc.cpu().numpy()
    err = np.abs(got - want).max() / np.abs(want).max()
    print("prolong", offset, periodic, f"{err:.2e}", flush=True)
    assert err <= 1e-13
    np.testing.assert_array_equal(got[act], child[act])


@pytest.mark.parametrize("ghost,offset,periodic", [(1, 0, True), (2, 0, True), (7, 0, True), (2, 6, False),
                                                   (7, 6, False)])
def test_prolongation_ghost_widths(J, ghost, offset, periodic):
    """Guard widths 1 and 2 take the column kernel for the z slabs; 7 needs
    18 parent z nodes per column (more than its sixteen lanes) and takes the
    point kernel."""
    from jax_bssn_amd import _lib
    lib = _lib.load()
    rng = np.random.default_rng(100 * ghost + offset)
    spec = OF.FMRPatchSpec((2, 3, 2), (6, 6, 7), ghost_width=ghost)
    pshape = tuple(n + 2 * offset for n in (12, 11, 13))
    cshape = tuple(n + 2 * ghost for n in OF.fine_active_shape(spec))
    parent = rng.standard_normal((21,) + pshape)
    child = rng.standard_normal((21,) + cshape)
    mats = OF.interpolation_matrices(pshape, spec, offset, periodic)
    act = (slice(None),) + OF.fine_active_slice(spec)
    want = OF.prolongate(parent, mats)
    want[act] = child[act]
    gp = torch.tensor(parent, device="cuda")
    gc = torch.tensor(child, device="cuda")
    print("prolong ghost", ghost, offset, periodic, pshape, cshape, flush=True)
    _lib.check(lib.bssn_fmr_prolong(_grid(pshape), offset, _i3((int(periodic),) * 3), _p(gp), _grid(cshape),
                                    _i3(spec.coarse_lo), ghost, _p(gc), None))
    got = gc.cpu().numpy()
    err = np.abs(got - want).max() / np.abs(want).max()
    print("prolong ghost", ghost, f"{err:.2e}", flush=True)
    assert err <= 1e-13
    np.testing.assert_array_equal(got[act], child[act])


def test_restriction_kernel_matches_oracle(J):
    from jax_bssn_amd import _lib
    lib = _lib.load()
    rng = np.random.default_rng(4)
    spec = OF.FMRPatchSpec((2, 3, 2), (6, 6, 7))
    pshape = (20, 19, 21)
    cshape = tuple(n + 8 for n in OF.fine_active_shape(spec))
    parent = rng.standard_normal((21,) + pshape)
    child = rng.standard_normal((21,) + cshape)
    want = parent.copy()
    cov = (slice(None),) + tuple(slice(4 + l, 4 + h + 1) for l, h in zip(spec.coarse_lo, spec.coarse_hi))
    want[cov] = child[(slice(None),) + OF.fine_active_slice(spec)][:, ::2, ::2, ::2]
    gp = torch.tensor(parent, device="cuda")
    gc = torch.tensor(child, device="cuda")
    _lib.check(lib.bssn_fmr_restrict(_grid(cshape), _i3(spec.coarse_lo), 4, _p(gc), _grid(pshape), 4, _p(gp),
                                     None))
    np.testing.assert_array_equal(gp.cpu().numpy(), want)


def test_transfer_argument_validation(J):
    from jax_bssn_amd import _lib
    lib = _lib.load()
    buf = ctypes.c_void_p(0x1000)
    # child allocation of even active extent
    with pytest.raises(ValueError, match="odd"):
        _lib.check(lib.bssn_fmr_prolong(_grid((12, 12, 12)), 0, _i3((1, 1, 1)), buf, _grid((16, 17, 17)),
                                        _i3((2, 2, 2)), 4, ctypes.c_void_p(0x2000), None))
    # stencil leaves a non-periodic parent
    with pytest.raises(ValueError, match="exceeds"):
        _lib.check(lib.bssn_fmr_prolong(_grid((12, 12, 12)), 0, _i3((0, 0, 0)), buf, _grid((17, 17, 17)),
                                        _i3((1, 2, 2)), 4, ctypes.c_void_p(0x2000), None))


def smooth_fmr_state(X, Y, Z):
    """Nonseparable gauge wave with a shift of both signs (the reference's
    tests/FMR/test_tiled_fmr.py smooth_state)."""
    X, Y, Z = np.broadcast_arrays(X, Y, Z)
    phase = X + 0.03 * np.sin(2 * np.pi * Y) * np.cos(2 * np.pi * Z)
    v = OI.gauge_wave_analytic_state(phase, Y, Z, amplitude=0.01)
    shift = np.stack((0.1 + 0.01 * np.sin(2 * np.pi * Y), -0.05 + 0.005 * np.cos(2 * np.pi * Z),
                      0.025 + 0.003 * np.sin(2 * np.pi * X)))
    return v._replace(shift=shift)


def chain(root_shape, params, specs, initial):
    """Root on x_min + i dx, children on their padded vertex allocations."""
    X, Y, Z = OB.grid_coordinates(root_shape, params)
    states = [initial(X, Y, Z)]
    parameters = [params]
    origin = (params.x_min, params.y_min, params.z_min)
    for spec in specs:
        coords = OF.fine_coordinates(spec, parameters[-1].dx, origin)
        dx = parameters[-1].dx / 2
        origin = tuple(o + lo * parameters[-1].dx for o, lo in zip(origin, spec.coarse_lo))
        states.append(initial(*coords))
        parameters.append(params._replace(dx=dx, x_min=origin[0] - 4 * dx, y_min=origin[1] - 4 * dx,
                                          z_min=origin[2] - 4 * dx))
    return states, parameters


@pytest.mark.parametrize("levels,steps,use_mad", [(2, 2, True), (3, 1, True), (2, 1, False)])
def test_fmr_steps_match_oracle(J, levels, steps, use_mad):
    from jax_bssn_amd import fmr as F
    specs = [OF.FMRPatchSpec((4, 2, 2), (12, 5, 5)), OF.FMRPatchSpec((4, 2, 2), (12, 4, 4))][:levels - 1]
    params = OB.BSSNParameters(eta=0.4, kappa=0.0, g=0.75, nu=0.05, dx=1 / 16, dt=0.25 / 64, zero_shift=0,
                               gauge=1)
    states, parameters = chain((16, 8, 8), params, specs, smooth_fmr_state)
    h = OF.FMRHierarchySpec(tuple(specs), use_mad=use_mad)
    ref = OF.fill_hierarchy(tuple(states), h)
    for _ in range(steps):
        ref = OF.fmr_rk4_step(ref, parameters, h)
    hj = F.FMRHierarchySpec(tuple(F.FMRPatchSpec(s.coarse_lo, s.coarse_hi) for s in specs), use_mad=use_mad)
    pj = [J.BSSNParameters(**p._asdict()) for p in parameters]
    ev = F.FMREvolver(states, pj, hj)
    ev.step(steps)
    got = ev.variables(like=np.zeros(1))
    for lv, (a, b) in enumerate(zip(got, ref)):
        errs = rel_errors(to_np(a), b)
        print("fmr", levels, steps, use_mad, "level", lv, " ".join(f"{k}={v:.1e}" for k, v in errs.items()),
              flush=True)
        assert max(errs.values()) <= STEP_TOL, (lv, errs)


def test_fmr_rk4_step_api(J):
    from jax_bssn_amd import fmr as F
    spec = OF.FMRPatchSpec((4, 2, 2), (12, 5, 5))
    params = OB.BSSNParameters(eta=0.0, g=0.0, nu=0.0, dx=1 / 16, dt=0.25 / 64, zero_shift=1, gauge=0)
    states, parameters = chain((16, 8, 8), params, [spec], smooth_fmr_state)
    h = OF.FMRHierarchySpec((spec,))
    ref = OF.fmr_rk4_step(tuple(states), parameters, h)
    got = F.fmr_rk4_step([J.BSSNVariables(*s) for s in states], [J.BSSNParameters(**p._asdict()) for p in parameters],
                         F.FMRHierarchySpec((F.FMRPatchSpec(spec.coarse_lo, spec.coarse_hi),)))
    for a, b in zip(got, ref):
        assert max(rel_errors(to_np(a), b).values()) <= STEP_TOL


# ------------------------------------------------------------ known answers

AMPLITUDE = 0.001
FINAL_TIME = 0.25


def _wave(X, Y, Z):
    X, Y, Z = np.broadcast_arrays(X, Y, Z)
    return OI.gauge_wave_analytic_state(X, Y, Z, amplitude=AMPLITUDE)


def _evolve_wave(J, n, steps, fixed_geometry):
    from jax_bssn_amd import fmr as F
    dx = 1 / n
    if fixed_geometry:
        shape = (n, n // 2, n // 2)
        spec = OF.FMRPatchSpec((n // 4, n // 8, n // 8), (3 * n // 4, 3 * n // 8, 3 * n // 8))
    else:
        shape = (n, 8, 8)
        spec = OF.FMRPatchSpec((n // 4, 2, 2), (3 * n // 4, 5, 5))
    params = OB.BSSNParameters(eta=0.0, kappa=0.0, g=0.0, nu=0.0, dx=dx, dt=FINAL_TIME / steps, zero_shift=1,
                               gauge=0)
    states, parameters = chain(shape, params, [spec], _wave)
    ev = F.FMREvolver(states, [J.BSSNParameters(**p._asdict()) for p in parameters],
                      F.FMRHierarchySpec((F.FMRPatchSpec(spec.coarse_lo, spec.coarse_hi),)))
    ev.step(steps)
    root = ev.u[0][17].cpu().numpy()
    fine = ev.u[1][17].cpu().numpy()[4:-4, 4:-4, 4:-4]
    uncovered = np.ones(shape, dtype=bool)
    uncovered[tuple(slice(lo, hi + 1) for lo, hi in zip(spec.coarse_lo, spec.coarse_hi))] = False
    root_x = dx * np.arange(n)[:, None, None]
    fine_x = OF.fine_coordinates(spec, dx)[0][4:-4, 4:-4, 4:-4]
    exact = tuple(np.sqrt(1 - AMPLITUDE * np.sin(2 * np.pi * (x - FINAL_TIME))) for x in (root_x, fine_x))
    return (root, fine), exact, uncovered, dx


def _composite_norm(fields, uncovered, dx):
    root, fine = fields
    volume = uncovered.sum() * dx ** 3 + fine.size * (dx / 2) ** 3
    square = np.where(uncovered, root ** 2, 0).sum() * dx ** 3 + (fine ** 2).sum() * (dx / 2) ** 3
    return math.sqrt(square / volume)


@pytest.mark.parametrize("n,published", [(16, 1.116773e-5), (24, 1.140085e-6)])
def test_published_spatial_errors(J, n, published):
    actual, exact, uncovered, dx = _evolve_wave(J, n, 256, fixed_geometry=True)
    err = _composite_norm(tuple((a - b) / AMPLITUDE for a, b in zip(actual, exact)), uncovered, dx)
    print("FMR spatial gauge-wave error", n, err, "published", published, flush=True)
    assert abs(err - published) <= 1e-5 * published


def test_published_temporal_differences(J):
    sols = []
    for steps in (20, 40, 80):
        actual, _, uncovered, dx = _evolve_wave(J, 24, steps, fixed_geometry=False)
        sols.append(actual)
    diffs = [_composite_norm(tuple((a - b) / AMPLITUDE for a, b in zip(c, f)), uncovered, dx)
             for c, f in zip(sols, sols[1:])]
    print("FMR temporal differences", diffs, "published", (2.8777354e-7, 4.1344261e-8), flush=True)
    for got, want in zip(diffs, (2.8777354e-7, 4.1344261e-8)):
        assert abs(got - want) <= 1e-5 * want
