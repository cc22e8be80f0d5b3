"""Pins of the oracle's RHS, projections, constraints and RK4 step.

What fixes each part:
* tensor algebra: numpy.linalg inverse / determinant;
* Sommerfeld: the r^2 known answer and flat-space stationarity of
  reference tests/BC/test_sommerfeld_boundaries.py;
* individual RHS terms: the closed-form constructions of reference
  tests/test_shift_evolution.py (flat metric, analytic shift / lapse);
* the Ricci / Hessian / constraint assembly: the Hamiltonian and momentum
  constraints of exact vacuum data (harmonic gauge wave, Brill-Lindquist
  puncture) vanish at fourth order, and the product-rule derivative of the
  momentum constraint converges to the derivative of the constraint;
* the full step: gauge-wave convergence (reference tests/test_gauge_wave.py)
  and det = 1, tr A = 0 after a step (tests/test_parallel.py:172-174).
"""

import numpy as np
import pytest

from oracle import bssn as B
from oracle import initial_data as I
from oracle.derivatives import SOMMERFELD_BC, diff1_field, diff1_upwind_field, diff2_field


def _random_spd(shape, rng, amp=0.2):
    g = np.zeros((3, 3) + shape)
    for i in range(3):
        g[i, i] = 1.0
    for i in range(3):
        for j in range(i, 3):
            e = amp * rng.uniform(-1, 1, shape)
            g[i, j] += e
            if i != j:
                g[j, i] += e
    return g


def test_inverse_and_determinant_against_numpy():
    rng = np.random.default_rng(0)
    g = _random_spd((4, 5, 6), rng)
    inv = B.invert_3x3_metric(g)
    gm = np.moveaxis(g.reshape(3, 3, -1), 2, 0)
    ref = np.moveaxis(np.linalg.inv(gm), 0, 2).reshape(g.shape)
    np.testing.assert_allclose(inv, ref, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(B.determinant_3x3_metric(g),
                               np.linalg.det(gm).reshape(4, 5, 6), rtol=1e-13)


def test_projection_gives_unit_determinant_and_trace_free_A():
    rng = np.random.default_rng(1)
    shape = (5, 4, 3)
    g = _random_spd(shape, rng)
    A = _random_spd(shape, rng, 0.5) - np.eye(3)[:, :, None, None, None]
    z = np.zeros(shape)
    v = B.BSSNVariables(g, z + 1, A, z, np.zeros((3,) + shape), z + 1, np.zeros((3,) + shape))
    p = B.enforce_algebraic_constraints(v)
    np.testing.assert_allclose(B.determinant_3x3_metric(p.conformal_metric), 1.0, atol=3e-15)
    np.testing.assert_allclose(B.compute_trace_A_violation(p), 0.0, atol=3e-14)


# ---------------------------------------------------------------- Sommerfeld

def _somm_setup():
    shape = (11, 9, 13)
    dx = 0.25
    p = B.BSSNParameters(dx=dx, xl_bc=1, xr_bc=1, yl_bc=1, yr_bc=1, zl_bc=1, zr_bc=1,
                         x_min=-5 * dx, y_min=-4 * dx, z_min=-6 * dx)
    X, Y, Z = B.grid_coordinates(shape, p)
    return shape, p, X, Y, Z


def _flat(shape):
    z = np.zeros(shape)
    g = np.broadcast_to(np.eye(3)[:, :, None, None, None], (3, 3) + shape).copy()
    return B.BSSNVariables(g, z + 1, np.zeros((3, 3) + shape), z.copy(),
                           np.zeros((3,) + shape), z + 1, np.zeros((3,) + shape))


def test_radial_derivative_of_r_squared_and_face_replacement():
    shape, p, X, Y, Z = _somm_setup()
    f = X ** 2 + Y ** 2 + Z ** 2 + 0 * X * Y * Z
    r = np.sqrt(f)
    np.testing.assert_allclose(B.radial_derivative(f, p), 2 * r, atol=2e-14)
    rhs = 7.0 * np.ones(shape)
    out = B.sommerfeld(f, rhs, 0.0, p)
    mask = B.sommerfeld_boundary_mask(shape, p)
    np.testing.assert_allclose(out[mask], -3.0 * r[mask], atol=3e-14)
    np.testing.assert_array_equal(out[~mask], rhs[~mask])


def test_mask_is_union_of_active_faces():
    shape, p, *_ = _somm_setup()
    p = p._replace(xl_bc=1, xr_bc=0, yl_bc=0, yr_bc=1, zl_bc=0, zr_bc=0)
    m = B.sommerfeld_boundary_mask(shape, p)
    e = np.zeros(shape, bool)
    e[0] = True
    e[:, -1] = True
    np.testing.assert_array_equal(m, e)


def test_assembled_rhs_keeps_flat_space_stationary():
    shape, p, *_ = _somm_setup()
    rhs = B.compute_bssn_rhs(_flat(shape), p._replace(nu=0.0))
    for f in rhs:
        assert np.all(np.isfinite(f))
        assert np.abs(f).max() < 1e-12


# ------------------------------------------------- term-level (shift tests)

N_TERM = 32


@pytest.fixture(scope="module")
def term_grid():
    x = np.linspace(-np.pi, np.pi, N_TERM, endpoint=False)
    dx = x[1] - x[0]
    X, Y, Z = np.meshgrid(x, x, x, indexing='ij')
    p = B.BSSNParameters(dx=dx, dt=0.01, nu=0.0, g=0.7, eta=0.2)
    return X, Y, Z, dx, p


def _term_state(X, shift=None, lapse=None, K=None, Gam=None):
    v = _flat(X.shape)
    return v._replace(shift=v.shift if shift is None else shift,
                      lapse=v.lapse if lapse is None else lapse,
                      trace_K=v.trace_K if K is None else K,
                      conformal_connection=v.conformal_connection if Gam is None else Gam)


def _beta(X, Y, Z):
    return np.stack([0.1 * np.sin(X), 0.2 * np.sin(Y), 0.3 * np.sin(Z)])


def _adv(f, beta, dx):
    lead = f.ndim - 3
    return sum(beta[d] * diff1_upwind_field(f, beta[d], lead + d, dx) for d in range(3))


def test_metric_rhs_is_weighted_lie_derivative(term_grid):
    X, Y, Z, dx, p = term_grid
    beta = _beta(X, Y, Z)
    r = B.compute_bssn_rhs(_term_state(X, shift=beta), p)
    db = np.stack([np.stack([diff1_field(beta[i], d, dx) for d in range(3)]) for i in range(3)])
    div = db[0, 0] + db[1, 1] + db[2, 2]
    exp = np.zeros((3, 3) + X.shape)
    for i in range(3):
        for j in range(3):
            exp[i, j] = db[i, j] + db[j, i] - (2 / 3) * (i == j) * div
    np.testing.assert_allclose(r.conformal_metric, exp, atol=2e-5)
    # analytic: d_j beta^i is diagonal, cos profiles
    np.testing.assert_allclose(db[0, 0], 0.1 * np.cos(X), atol=6e-5)


def test_connection_rhs_has_second_shift_derivatives(term_grid):
    X, Y, Z, dx, p = term_grid
    beta = _beta(X, Y, Z)
    r = B.compute_bssn_rhs(_term_state(X, shift=beta), p)
    amp = (0.1, 0.2, 0.3)
    axes = (X, Y, Z)
    # flat space: dt Gamma^i = lap beta^i + 1/3 d_i div beta  (exact: -4/3 a_i sin)
    exact = np.stack([-(4 / 3) * amp[i] * np.sin(axes[i]) for i in range(3)])
    exact = exact + _adv(np.zeros((3,) + X.shape), beta, dx)
    np.testing.assert_allclose(r.conformal_connection, exact, atol=3e-5)


@pytest.mark.parametrize("gauge", [0, 1])
def test_lapse_rhs_slicing_and_advection(term_grid, gauge):
    X, Y, Z, dx, p = term_grid
    beta = _beta(X, Y, Z)
    alpha = 1.0 + 0.05 * np.cos(X + Y)
    K = 0.02 * np.sin(X + Z)
    r = B.compute_bssn_rhs(_term_state(X, shift=beta, lapse=alpha, K=K), p._replace(gauge=gauge))
    slicing = -alpha ** 2 * K if gauge == 0 else -2.0 * alpha * K
    np.testing.assert_allclose(r.lapse, slicing + _adv(alpha, beta, dx), atol=2e-5)
    # the advection itself: beta . grad alpha in closed form
    grad = np.stack([-0.05 * np.sin(X + Y), -0.05 * np.sin(X + Y), 0 * X])
    np.testing.assert_allclose(r.lapse - slicing, (beta * grad).sum(0), atol=2e-4)


def test_shift_rhs_gamma_driver_and_zero_shift(term_grid):
    X, Y, Z, dx, p = term_grid
    beta = _beta(X, Y, Z)
    Gam = np.stack([0.01 * np.cos(X), 0.02 * np.cos(Y), 0.03 * np.cos(Z)])
    v = _term_state(X, shift=beta, Gam=Gam)
    r = B.compute_bssn_rhs(v, p)
    np.testing.assert_allclose(r.shift, p.g * Gam + _adv(beta, beta, dx) - p.eta * beta,
                               atol=2e-5)
    out = B.rk4_step(v, p._replace(zero_shift=1, dt=0.001))
    np.testing.assert_array_equal(out.shift, beta)
    out = B.rk4_step(_term_state(X, Gam=Gam), p._replace(zero_shift=0, dt=0.001))
    assert np.abs(out.shift).max() > 0.0


# ------------------------------------------- constraints of exact vacuum data

def _gauge_wave(n, t=0.0):
    X, Y, Z = I.periodic_coordinate_arrays(n)
    return I.gauge_wave_analytic_state(X, Y, Z, t=t), B.BSSNParameters(dx=1.0 / n)


def test_gauge_wave_constraints_vanish_at_fourth_order():
    errs_h, errs_m, errs_g = [], [], []
    for n in (16, 32):
        v, p = _gauge_wave(n)
        c = B.compute_all_constraints(v, p)
        errs_h.append(np.sqrt(np.mean(c.hamiltonian ** 2)))
        errs_m.append(np.sqrt(np.mean(c.momentum ** 2)))
        errs_g.append(np.sqrt(np.mean(c.gamma_condition ** 2)))
    for e in (errs_h, errs_m, errs_g):
        assert np.log2(e[0] / e[1]) > 3.8, e


def test_puncture_hamiltonian_constraint_vanishes_at_fourth_order():
    """Brill-Lindquist data is vacuum (H = 0); box [1, 3]^3 with Sommerfeld
    closures on every face (the puncture at the origin is outside)."""
    errs = []
    for n in (16, 32):
        dx = 2.0 / (n - 1)
        p = B.BSSNParameters(dx=dx, xl_bc=1, xr_bc=1, yl_bc=1, yr_bc=1, zl_bc=1, zr_bc=1,
                             x_min=1.0, y_min=1.0, z_min=1.0)
        X, Y, Z = B.grid_coordinates((n, n, n), p)
        X, Y, Z = np.broadcast_arrays(X, Y, Z)
        v = I.puncture_state(X, Y, Z, masses=(1.0,))
        H = B.compute_hamiltonian_constraint(v, p)
        errs.append(np.sqrt(np.mean(H ** 2)))
    assert np.log2(errs[0] / errs[1]) > 3.5, errs


def _smooth_state(n, seed=5):
    """A non-solution smooth periodic state exercising every term."""
    rng = np.random.default_rng(seed)
    x = np.arange(n) / n
    X, Y, Z = np.meshgrid(x, x, x, indexing='ij')
    tp = 2 * np.pi

    def mode(a):
        k = rng.integers(1, 3, 3)
        ph = rng.uniform(0, tp, 3)
        return a * np.sin(tp * k[0] * X + ph[0]) * np.cos(tp * k[1] * Y + ph[1]) \
            * np.sin(tp * k[2] * Z + ph[2])
    g = np.zeros((3, 3, n, n, n))
    for i in range(3):
        for j in range(i, 3):
            g[i, j] = (1.0 if i == j else 0.0) + mode(0.05)
            g[j, i] = g[i, j]
    A = np.zeros_like(g)
    for i in range(3):
        for j in range(i, 3):
            A[i, j] = mode(0.1)
            A[j, i] = A[i, j]
    v = B.BSSNVariables(g, 1.0 + mode(0.1), A, mode(0.1),
                        np.stack([mode(0.1) for _ in range(3)]), 1.0 + mode(0.1),
                        np.stack([mode(0.1) for _ in range(3)]))
    return B.enforce_algebraic_constraints(v), B.BSSNParameters(dx=1.0 / n)


def test_momentum_constraint_product_rule_derivative_converges():
    """dM (product-rule, fed to the kappa term) -> d_l M_i at fourth order."""
    diffs = []
    for n in (16, 32):
        v, p = _smooth_state(n)
        M, dM = B.compute_momentum_constraint_and_derivative(v, p)
        dnum = np.stack([np.stack([diff1_field(M[i], l, p.dx) for l in range(3)])
                         for i in range(3)])
        diffs.append(np.sqrt(np.mean((dM - dnum) ** 2)) / np.sqrt(np.mean(dnum ** 2)))
    assert np.log2(diffs[0] / diffs[1]) > 3.5, diffs


def test_kappa_term_is_the_symmetrised_covariant_derivative():
    v, p = _smooth_state(12, seed=9)
    r0 = B.compute_bssn_rhs(v, p._replace(kappa=0.0))
    r1 = B.compute_bssn_rhs(v, p._replace(kappa=0.3))
    M, dM = B.compute_momentum_constraint_and_derivative(v, p)
    inv = B.invert_3x3_metric(v.conformal_metric)
    c2 = B.christoffel_symbols_second_kind(
        inv, np.stack([diff1_field(v.conformal_metric, d + 2, p.dx) for d in range(3)]))
    DjMi = dM - np.einsum('kij...,k...->ij...', c2, M)
    exp = 0.15 * v.lapse * (DjMi + np.swapaxes(DjMi, 0, 1))
    np.testing.assert_allclose(r1.traceless_K - r0.traceless_K, exp, atol=1e-12)
    for a, b in zip(r0[:2] + r0[3:], r1[:2] + r1[3:]):
        np.testing.assert_array_equal(a, b)


# ------------------------------------------------------------- full RK4 step

def test_step_keeps_algebraic_constraints():
    v, p = _smooth_state(12)
    out = B.rk4_step(v, p._replace(dt=0.01))
    np.testing.assert_allclose(B.compute_det_gamma_violation(out), 0.0, atol=3e-14)
    np.testing.assert_allclose(B.compute_trace_A_violation(out), 0.0, atol=3e-14)


@pytest.mark.slow
def test_gauge_wave_convergence_order():
    """tests/test_gauge_wave.py:67-138: 16 -> 32, dt = dx, T = 1, nu = 0."""
    errs = []
    for n in (16, 32):
        X, Y, Z = I.periodic_coordinate_arrays(n)
        dx = 1.0 / n
        v = I.periodic_gauge_wave_state(X, Y, Z, dx)
        p = B.BSSNParameters(eta=0.0, kappa=0.0, g=0.0, dx=dx, dt=dx, nu=0.0,
                             zero_shift=1, gauge=0)
        for _ in range(n):
            v = B.rk4_step(v, p)
        ex = I.gauge_wave_analytic_state(X, Y, Z, 1.0)
        errs.append([np.sqrt(np.mean((v.conformal_metric[i, i] - ex.conformal_metric[i, i]) ** 2))
                     for i in range(3)]
                    + [np.sqrt(np.mean((v.conformal_factor - ex.conformal_factor) ** 2))])
    orders = np.log2(np.array(errs[0]) / np.array(errs[1]))
    assert np.all(orders > 3.8), orders


def test_matter_sources_enter_with_reference_coefficients():
    """Flat space (vacuum RHS = 0): dt K = 4 pi alpha (rho + S), dt A_ij =
    -8 pi alpha (S_ij - gamma_ij S / 3), dt Gamma^i = -16 pi alpha S_i,
    H -> H - 16 pi rho, M_i -> M_i - 8 pi S_i."""
    shape = (8, 9, 10)
    rng = np.random.default_rng(4)
    v = _flat(shape)
    rho = rng.uniform(0.1, 0.2, shape)
    S = rng.uniform(-0.1, 0.1, (3,) + shape)
    Sij = rng.uniform(-0.1, 0.1, (3, 3) + shape)
    Sij = 0.5 * (Sij + np.swapaxes(Sij, 0, 1))
    p = B.BSSNParameters(dx=0.1, nu=0.0)
    r = B.compute_bssn_rhs_with_matter(v, p, rho, S, Sij)
    trS = Sij[0, 0] + Sij[1, 1] + Sij[2, 2]
    np.testing.assert_allclose(r.trace_K, 4 * np.pi * (rho + trS), atol=1e-14)
    expA = -8 * np.pi * (Sij - np.eye(3)[:, :, None, None, None] * trS / 3)
    np.testing.assert_allclose(r.traceless_K, expA, atol=1e-14)
    np.testing.assert_allclose(r.conformal_connection, -16 * np.pi * S, atol=1e-14)
    for f in (r.conformal_metric, r.conformal_factor, r.lapse, r.shift):
        assert np.abs(f).max() < 1e-14
    c = B.compute_all_constraints_with_matter(v, p, rho, S)
    np.testing.assert_allclose(c.hamiltonian, -16 * np.pi * rho, atol=1e-13)
    np.testing.assert_allclose(c.momentum, -8 * np.pi * S, atol=1e-13)


def test_complex_step_through_the_oracle_is_its_derivative():
    """The oracle runs on complex fields (the upwind side chosen by the real
    part of the shift), so Im F(u + i h du) / h is the exact derivative of
    its RHS and of its RK4 step: it must match real central differences of
    the same maps to their truncation error.  The adjoint tests use this
    derivative as the stand-in for jax.grad of the reference."""
    import warnings
    from tests.parity_utils import smooth_state
    v = smooth_state((10, 12, 14), seed=5)
    p = B.BSSNParameters(eta=0.3, g=0.75, nu=0.05, dx=0.1, dt=0.025, zero_shift=0, gauge=1, xl_bc=1, xr_bc=1,
                         zl_bc=1, x_min=-0.45, y_min=-0.55, z_min=-0.65)
    rng = np.random.default_rng(3)
    sym = lambda t: 0.5 * (t + np.swapaxes(t, 0, 1))
    dv = B.BSSNVariables(*[1e-2 * rng.standard_normal(np.shape(x)) for x in v])
    dv = dv._replace(conformal_metric=sym(dv.conformal_metric), traceless_K=sym(dv.traceless_K))
    h, e = 1e-30, 1e-6
    for fn in (lambda u: B.compute_bssn_rhs(u, p), lambda u: B.rk4_step(u, p)):
        with warnings.catch_warnings():
            warnings.simplefilter("error", np.exceptions.ComplexWarning)  # no silent cast to real
            cs = fn(B.BSSNVariables(*[np.asarray(a, dtype=complex) + 1j * h * d for a, d in zip(v, dv)]))
        plus = fn(B.BSSNVariables(*[a + e * d for a, d in zip(v, dv)]))
        minus = fn(B.BSSNVariables(*[a - e * d for a, d in zip(v, dv)]))
        for c, a, b in zip(cs, plus, minus):
            fd = (a - b) / (2 * e)
            assert np.abs(np.imag(c) / h - fd).max() <= 1e-7 * np.abs(fd).max()
            assert np.abs(np.real(c) - 0.5 * (a + b)).max() <= 1e-9 * max(np.abs(a).max(), 1.0)
