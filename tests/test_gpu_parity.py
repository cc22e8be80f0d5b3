"""GPU parity: the HIP path against the FP64 oracle on the same seeded inputs.

Bars (stated here, used below):
* one RHS evaluation, FP64: per-field max-norm relative difference
  <= RHS_TOL.  The RHS contains second differences (roundoff ~ eps |u| / dx^2)
  evaluated with a different summation order than the oracle, so its bar is
  looser than the state bar;
* state after each RK4 step, FP64: <= 1e-12;
* FP32-RHS: <= F32_RHS_TOL per RHS, <= 1e-6 per step.
"""

import numpy as np
import pytest
import torch

from oracle import bssn as OB
from oracle import initial_data as OI
from tests.parity_utils import rel_errors, smooth_state, to_np

pytestmark = pytest.mark.gpu

RHS_TOL = 1e-10
STEP_TOL = 1e-12
F32_RHS_TOL = 5e-4


@pytest.fixture(scope="module")
def bssn():
    import jax_bssn_amd as J
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    return J


def _params(J, p):
    return J.BSSNParameters(**p._asdict())


def _log(name, errs):
    print(name, " ".join(f"{k}={v:.2e}" for k, v in errs.items()), flush=True)


def gauge_case(n=24):
    X, Y, Z = OI.periodic_coordinate_arrays(n)
    v = OI.gauge_wave_analytic_state(X, Y, Z)
    p = OB.BSSNParameters(eta=0.0, kappa=0.0, g=0.0, nu=0.1, dx=1.0 / n, dt=0.25 / n,
                          zero_shift=1, gauge=0)
    return v, p


def shift_case(shape=(16, 16, 16), seed=5):
    v = smooth_state(shape, seed=seed)
    p = OB.BSSNParameters(eta=0.3, g=0.75, nu=0.05, dx=1.0 / shape[0], dt=0.2 / shape[0],
                          zero_shift=0, gauge=1)
    return v, p


def sommerfeld_case(shape=(16, 17, 18), seed=7, faces=(1, 1, 1, 1, 1, 1)):
    v = smooth_state(shape, seed=seed)
    dx = 0.1
    p = OB.BSSNParameters(eta=2.0, g=0.75, nu=0.1, dx=dx, dt=0.25 * dx, zero_shift=0, gauge=1,
                          xl_bc=faces[0], xr_bc=faces[1], yl_bc=faces[2], yr_bc=faces[3],
                          zl_bc=faces[4], zr_bc=faces[5],
                          x_min=-0.75, y_min=-0.8, z_min=-0.83)
    return v, p


CASES = {
    "gauge_wave": lambda: gauge_case(24),
    "shift_1plog": lambda: shift_case(),
    "noncubic_shift": lambda: shift_case((12, 16, 20), seed=11),
    "sommerfeld_all": lambda: sommerfeld_case(),
    "sommerfeld_mixed": lambda: sommerfeld_case((18, 16, 14), seed=9, faces=(1, 0, 0, 1, 1, 1)),
    "mad_periodic": lambda: (lambda v, p: (v, p._replace(mad_q=1.0 / 16.0)))(*shift_case(seed=3)),
    "mad_mixed_sommerfeld": lambda: (lambda v, p: (v, p._replace(mad_q=0.25)))(
        *sommerfeld_case((16, 16, 16), seed=4, faces=(0, 0, 1, 1, 0, 0))),
    "zero_shift_harmonic": lambda: (lambda v, p: (v, p._replace(zero_shift=1, gauge=0)))(*shift_case(seed=8)),
    # shapes that tile exactly: closures evaluated inside the LDS tiles
    "sommerfeld_aligned": lambda: sommerfeld_case((16, 16, 32), seed=12),
    # 24^3: face tiles of every kind (x, y and z slabs with one closure axis,
    # edges and corners) each take their own kernel
    "sommerfeld_aligned_all_slabs": lambda: sommerfeld_case((24, 24, 24), seed=15),
    "kappa_periodic": lambda: (lambda v, p: (v, p._replace(kappa=0.3)))(*shift_case(seed=13)),
    "kappa_sommerfeld_aligned": lambda: (lambda v, p: (v, p._replace(kappa=0.2)))(
        *sommerfeld_case((16, 16, 32), seed=14)),
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_rhs_f64_matches_oracle(bssn, case):
    v, p = CASES[case]()
    print(f"[rhs f64] case={case} shape={v.conformal_factor.shape}", flush=True)
    ref = OB.compute_bssn_rhs(v, p)
    got = to_np(bssn.compute_bssn_rhs(v, _params(bssn, p)))
    errs = rel_errors(got, ref)
    _log(case, errs)
    assert max(errs.values()) <= RHS_TOL, errs


@pytest.mark.parametrize("case", ["gauge_wave", "shift_1plog", "sommerfeld_all"])
def test_rhs_f32_within_stated_error(bssn, case):
    v, p = CASES[case]()
    print(f"[rhs f32] case={case}", flush=True)
    ref = OB.compute_bssn_rhs(v, p)
    got = to_np(bssn.compute_bssn_rhs(v, _params(bssn, p), precision="f32rhs"))
    errs = rel_errors(got, ref)
    _log(case, errs)
    assert max(errs.values()) <= F32_RHS_TOL, errs


def test_projection_matches_oracle(bssn):
    v = smooth_state((10, 11, 12), seed=2)
    g = v.conformal_metric * 1.1
    v = v._replace(conformal_metric=g, traceless_K=v.traceless_K + 0.05 * g)
    ref = OB.enforce_algebraic_constraints(v)
    got = to_np(bssn.enforce_algebraic_constraints(v))
    errs = rel_errors(got, ref)
    assert max(errs.values()) <= 1e-14, errs


@pytest.mark.parametrize("case,steps", [("gauge_wave", 6), ("shift_1plog", 4), ("sommerfeld_all", 3),
                                        ("mad_periodic", 2), ("sommerfeld_aligned", 2),
                                        ("kappa_sommerfeld_aligned", 2), ("sommerfeld_aligned_all_slabs", 1)])
def test_rk4_steps_match_oracle_every_step(bssn, case, steps):
    v, p = CASES[case]()
    print(f"[rk4] case={case} steps={steps}", flush=True)
    ev = bssn.Evolver(bssn.pack(v), _params(bssn, p))
    ref = v
    for s in range(steps):
        ref = OB.rk4_step(ref, p)
        ev.step(1)
        got = to_np(ev.variables(like=np.zeros(1)))
        errs = rel_errors(got, ref)
        _log(f"{case} step {s + 1}", errs)
        assert max(errs.values()) <= STEP_TOL, (s, errs)


def test_rk4_step_api_equals_evolver_steps(bssn):
    v, p = CASES["shift_1plog"]()
    once = to_np(bssn.rk4_step(bssn.rk4_step(v, _params(bssn, p)), _params(bssn, p)))
    twice = to_np(bssn.evolve(v, _params(bssn, p), 2))
    # same kernels on the same buffers up to launch grouping: equal to rounding
    for a, b in zip(once, twice):
        np.testing.assert_allclose(a, b, rtol=1e-13, atol=1e-15)


def test_constraints_match_oracle(bssn):
    # periodic and aligned Sommerfeld grids take the tile kernels (lean and
    # face tiles), the others the point kernel
    for case in ("shift_1plog", "sommerfeld_all", "sommerfeld_aligned", "sommerfeld_aligned_all_slabs",
                 "sommerfeld_mixed", "mad_periodic"):
        v, p = CASES[case]()
        ref = OB.compute_all_constraints(v, p)
        got = bssn.compute_all_constraints(v, _params(bssn, p))
        for name in ref._fields:
            a = np.asarray(getattr(got, name))
            b = np.asarray(getattr(ref, name))
            scale = max(np.abs(b).max(), 1e-300)
            err = np.abs(a - b).max() / scale
            print(case, name, f"{err:.2e}", flush=True)
            assert err <= 1e-10 or np.abs(a - b).max() < 1e-13, (case, name, err)
        rn = OB.compute_constraint_norms(ref)
        gn = bssn.compute_constraint_norms(got)
        for k in rn:
            # norms of roundoff-level fields (det - 1, tr A ~ 1e-16) get an absolute floor
            assert abs(gn[k] - rn[k]) <= 1e-10 * abs(rn[k]) + 1e-14, (k, gn[k], rn[k])


@pytest.mark.parametrize("case", ["shift_1plog", "kappa_periodic", "sommerfeld_aligned"])
def test_rhs_with_matter_matches_oracle(bssn, case):
    from tests.test_host_algebra import matter_sources
    v, p = CASES[case]()
    rho, S, Sij = matter_sources(v.conformal_factor.shape)
    print(f"[matter] case={case}", flush=True)
    ref = OB.compute_bssn_rhs_with_matter(v, p, rho, S, Sij)
    got = to_np(bssn.compute_bssn_rhs_with_matter(v, _params(bssn, p), rho, S, Sij))
    errs = rel_errors(got, ref)
    _log(case, errs)
    assert max(errs.values()) <= RHS_TOL, errs
    cref = OB.compute_all_constraints_with_matter(v, p, rho, S)
    cgot = bssn.compute_all_constraints_with_matter(v, _params(bssn, p), rho, S)
    for name in ("hamiltonian", "momentum"):
        a, b = np.asarray(getattr(cgot, name)), np.asarray(getattr(cref, name))
        assert np.abs(a - b).max() <= 1e-10 * np.abs(b).max(), name


def test_f32rhs_puncture_face_tiles_stay_finite(bssn):
    """Regression: the face-tile kernel's mixed derivatives once read
    zero-weight taps outside the LDS box; harmless in FP64, NaN in FP32.
    The 128^3 puncture (tile-aligned Sommerfeld faces) exercises every face
    tile; the FP32-RHS RHS must be finite and within F32_RHS_TOL."""
    import bench
    from jax_bssn_amd.time_evolve import rhs_packed
    state, dx, x0 = bench.puncture_state(128)
    p = bench.puncture_params(bssn.BSSNParameters, dx, x0)
    r32 = rhs_packed(state, p, "f32rhs")
    r64 = rhs_packed(state, p, "f64")
    assert bool(torch.isfinite(r32).all())
    scale = r64.abs().reshape(21, -1).max(1).values
    err = (r32 - r64).abs().reshape(21, -1).max(1).values
    live = scale > 1e-12
    worst = float((err[live] / scale[live]).max())
    print("f32rhs puncture 128^3 rhs worst", f"{worst:.1e}", flush=True)
    assert worst <= F32_RHS_TOL
