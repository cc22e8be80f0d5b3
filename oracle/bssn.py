"""FP64 NumPy restatement of the reference BSSN right-hand side and RK4 step.

TEST INFRASTRUCTURE ONLY.  This module is the CPU oracle against which the HIP
kernels are checked.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import it; the product path never
does.

It follows, function by function, the reference files

* bssn/variables.py            (state / parameter containers)
* bssn/tensor_algebra.py       (Cholesky inverse, determinant, Christoffels)
* bssn/geometry.py             (W^2 R_ij, W^2 D_i D_j alpha)
* bssn/spatial_metric.py, bssn/extrinsic_curvature.py,
  bssn/conformal_connection.py, bssn/shift_and_lapse.py   (the 7 RHS)
* bssn/constraints.py          (momentum constraint and its derivative)
* evolution/boundaries.py      (Sommerfeld RHS replacement)
* evolution/time_evolve.py     (RHS assembly, projections, RK4)

The only liberty taken is that every derivative array a RHS evaluation needs
is computed once and reused (the reference recomputes several of them inside
each equation); the stencils and the algebra applied to them are the same.

Pins: see tests/test_oracle_bssn.py (flat space stationary with Sommerfeld
faces, the r^2 Sommerfeld known answer, term-level shift/lapse checks against
closed forms, gauge-wave accuracy and convergence order, det = 1 and
tr A = 0 after a step).  The bit-level roundoff of einsum orders is not
pinned (no reference-held vectors exist).
"""

from typing import NamedTuple

import numpy as np

from .derivatives import (
    SOMMERFELD_BC,
    diff1_physical,
    diff1_upwind_physical,
    diff2_physical,
    ko_dissipation,
)

W_FLOOR_VALUE = 1.0e-12  # reference bssn/geometry.py:19


class BSSNVariables(NamedTuple):
    """State container, same field order as reference bssn/variables.py:8-16."""
    conformal_metric: np.ndarray
    conformal_factor: np.ndarray
    traceless_K: np.ndarray
    trace_K: np.ndarray
    conformal_connection: np.ndarray
    lapse: np.ndarray
    shift: np.ndarray


class BSSNParameters(NamedTuple):
    """Parameters with the reference defaults (bssn/variables.py:19-46)."""
    eta: float = 2.0
    kappa: float = 0.0
    nu: float = 0.002
    g: float = 0.75
    dx: float = 0.1
    dt: float = 0.001
    zero_shift: int = 0
    gauge: int = 0
    xl_bc: int = 0
    xr_bc: int = 0
    yl_bc: int = 0
    yr_bc: int = 0
    zl_bc: int = 0
    zr_bc: int = 0
    x_min: float = 0.0
    y_min: float = 0.0
    z_min: float = 0.0
    mad_q: float = 1.0
    guard_cells: int = 0
    owned_shape: tuple = None
    owned_faces: tuple = (True, True, True, True, True, True)
    reduced_axes: tuple = None


# --------------------------------------------------------------------------
# Tensor algebra (reference bssn/tensor_algebra.py:20-193, 337-381)
# --------------------------------------------------------------------------

def invert_3x3_metric(metric):
    """Explicit Cholesky inverse, upper triangle read (tensor_algebra.py:20-71)."""
    g00, g01, g02 = metric[0, 0], metric[0, 1], metric[0, 2]
    g11, g12, g22 = metric[1, 1], metric[1, 2], metric[2, 2]
    L00 = np.sqrt(g00)
    L10 = g01 / L00
    L20 = g02 / L00
    L11 = np.sqrt(g11 - L10 * L10)
    L21 = (g12 - L20 * L10) / L11
    L22 = np.sqrt(g22 - L20 * L20 - L21 * L21)
    M00 = 1.0 / L00
    M10 = -L10 / (L00 * L11)
    M11 = 1.0 / L11
    M20 = (L10 * L21 - L20 * L11) / (L00 * L11 * L22)
    M21 = -L21 / (L11 * L22)
    M22 = 1.0 / L22
    inv00 = M00 * M00 + M10 * M10 + M20 * M20
    inv01 = M10 * M11 + M20 * M21
    inv02 = M20 * M22
    inv11 = M11 * M11 + M21 * M21
    inv12 = M21 * M22
    inv22 = M22 * M22
    return np.stack((np.stack((inv00, inv01, inv02)),
                     np.stack((inv01, inv11, inv12)),
                     np.stack((inv02, inv12, inv22))))


def determinant_3x3_metric(metric):
    """Pointwise 3x3 determinant (tensor_algebra.py:74-96, LU via numpy)."""
    g = np.moveaxis(metric.reshape(3, 3, -1), 2, 0)
    return np.linalg.det(g).reshape(metric.shape[2:])


def christoffel_symbols_first_kind(metric_derivatives):
    """Gamma_ijk = (d_k g_ij + d_j g_ik - d_i g_jk) / 2 (tensor_algebra.py:148-170)."""
    md = metric_derivatives
    return 0.5 * (np.einsum('kij...->ijk...', md) + np.einsum('jik...->ijk...', md)
                  - np.einsum('ijk...->ijk...', md))


def christoffel_symbols_second_kind(inverse_metric, metric_derivatives):
    """Gamma^i_jk = g^im Gamma_mjk (tensor_algebra.py:173-193)."""
    c1 = christoffel_symbols_first_kind(metric_derivatives)
    return np.einsum('im...,mjk...->ijk...', inverse_metric, c1)


def trace_tensor(tensor, inverse_metric):
    return np.einsum('ij...,ij...->...', inverse_metric, tensor)


def traceless_part(tensor, metric, inverse_metric):
    trace = trace_tensor(tensor, inverse_metric)
    return tensor - (1.0 / 3.0) * metric * trace


# --------------------------------------------------------------------------
# Derivative table of one RHS evaluation
# --------------------------------------------------------------------------

class _Derivatives:
    """Every derivative array one RHS evaluation uses, computed once.

    ``d1(x)[d]`` = diff1_physical along spatial axis d; ``d2(x)[m][n]`` =
    diff2_physical when m == n, else diff1_physical(d1(x)[m], n) -- the inner
    derivative on m and the outer on n, the composition order of the
    reference (geometry.py:50-56, 93-99, 143-148).
    """

    def __init__(self, params):
        self.params = params
        self._d1 = {}
        self._d2 = {}

    def d1(self, key, field):
        if key not in self._d1:
            lead = field.ndim - 3
            self._d1[key] = np.stack([diff1_physical(field, lead + d, self.params)
                                      for d in range(3)])
        return self._d1[key]

    def d2(self, key, field):
        if key not in self._d2:
            lead = field.ndim - 3
            first = self.d1(key, field)
            rows = []
            for m in range(3):
                row = []
                for n in range(3):
                    if m == n:
                        row.append(diff2_physical(field, lead + n, self.params))
                    else:
                        row.append(diff1_physical(first[m], lead + n, self.params))
                rows.append(np.stack(row))
            self._d2[key] = np.stack(rows)
        return self._d2[key]


def compute_shift_advection(field, shift, params):
    """beta^i D^upwind_i field (bssn/shift_and_lapse.py:43-62)."""
    lead = field.ndim - 3
    terms = [shift[d] * diff1_upwind_physical(field, shift[d], lead + d, params)
             for d in range(3)]
    return terms[0] + terms[1] + terms[2]


# --------------------------------------------------------------------------
# Geometry (reference bssn/geometry.py)
# --------------------------------------------------------------------------

def _w2_ricci(v, inv_gamma, D):
    """W^2 R_ij without inverse powers of W (geometry.py:21-119)."""
    gamma = v.conformal_metric
    W = v.conformal_factor
    dg = D.d1('gamma', gamma)                      # (d, i, j, ...)
    ddg = D.d2('gamma', gamma)                     # (m, n, i, j, ...)
    c1 = christoffel_symbols_first_kind(dg)
    c2 = christoffel_symbols_second_kind(inv_gamma, dg)
    term_1 = np.zeros_like(gamma)
    for m in range(3):
        for n in range(3):
            term_1 = term_1 - 0.5 * (inv_gamma[m, n] * ddg[m, n])
    dGam = D.d1('Gamma', v.conformal_connection)   # (d, i, ...)
    term_2 = (np.einsum('mi...,jm...->ij...', gamma, dGam)
              + np.einsum('mj...,im...->ij...', gamma, dGam)) / 2.0
    Gam = v.conformal_connection
    term_3 = (np.einsum('m...,ijm...->ij...', Gam, c1)
              + np.einsum('m...,jim...->ij...', Gam, c1)) / 2.0
    term_4 = (np.einsum('mn...,kmi...,jkn...->ij...', inv_gamma, c2, c1)
              + np.einsum('mn...,kmj...,ikn...->ij...', inv_gamma, c2, c1))
    term_5 = np.einsum('mn...,kim...,kjn...->ij...', inv_gamma, c2, c1)
    conformal_ricci = term_1 + term_2 + term_3 + term_4 + term_5
    dW = D.d1('W', W)
    ddW = D.d2('W', W)
    DiDjW = ddW - np.einsum('kij...,k...->ij...', c2, dW)
    lap_W = np.einsum('mn...,mn...->...', inv_gamma, DiDjW)
    grad_W2 = np.einsum('mn...,m...,n...->...', inv_gamma, dW, dW)
    w2_r_w = W * (DiDjW + gamma * lap_W) - 2.0 * gamma * grad_W2
    return W ** 2 * conformal_ricci + w2_r_w


def _w2_lapse_hessian(v, inv_gamma, D):
    """W^2 D_i D_j alpha (geometry.py:122-186)."""
    alpha = v.lapse
    W = v.conformal_factor
    gamma = v.conformal_metric
    da = D.d1('alpha', alpha)
    dda = D.d2('alpha', alpha)
    dW = D.d1('W', W)
    c2 = christoffel_symbols_second_kind(inv_gamma, D.d1('gamma', gamma))
    hess = dda - np.einsum('kij...,k...->ij...', c2, da)
    grad = (np.einsum('i...,j...->ij...', dW, da) + np.einsum('j...,i...->ij...', dW, da)
            - np.einsum('ij...,mn...,m...,n...->ij...', gamma, inv_gamma, dW, da))
    return W ** 2 * hess + W * grad


def compute_W2_ricci(v, params):
    return _w2_ricci(v, invert_3x3_metric(v.conformal_metric), _Derivatives(params))


def compute_W2_covariant_lapse_hessian(v, params):
    return _w2_lapse_hessian(v, invert_3x3_metric(v.conformal_metric),
                             _Derivatives(params))


# --------------------------------------------------------------------------
# Momentum constraint and its derivative (bssn/constraints.py:22-213)
# --------------------------------------------------------------------------

def _momentum_constraint(v, inv_gamma, D):
    K = v.trace_K
    A = v.traceless_K
    W_floor = np.maximum(v.conformal_factor, W_FLOOR_VALUE)
    dK = D.d1('K', K)
    dW = D.d1('W', v.conformal_factor)
    B = np.einsum('jk...,ik...->ij...', inv_gamma, A)      # A_i^j
    dB = D.d1('B', B)                                         # (k, i, j)
    dA = D.d1('A', A)
    first = np.einsum('jij...->i...', dB)
    second = -0.5 * np.einsum('jk...,ijk...->i...', inv_gamma, dA)
    third = -3 * np.einsum('ij...,j...->i...', B, dW) / W_floor
    fourth = -2.0 / 3.0 * dK
    return first + second + third + fourth


def _momentum_constraint_and_derivative(v, inv_gamma, D):
    A = v.traceless_K
    W = v.conformal_factor
    W_floor = np.maximum(W, W_FLOOR_VALUE)
    K = v.trace_K
    B = np.einsum('jk...,ik...->ij...', inv_gamma, A)
    B_over_W = B / W_floor
    M = _momentum_constraint(v, inv_gamma, D)
    dB = D.d1('B', B)
    dA = D.d1('A', A)
    d_inv = D.d1('inv_gamma', inv_gamma)
    dBW = D.d1('B_over_W', B_over_W)
    dW = D.d1('W', W)
    # D.d2(...)[a, b] is the outer derivative b of the inner derivative a;
    # the reference's d2X[l, inner] (outer l of inner ``inner``) is [inner, l].
    d2B = D.d2('B', B)
    d2A = D.d2('A', A)
    d2W = D.d2('W', W)
    d2K = D.d2('K', K)
    dM = np.zeros((3, 3) + K.shape, dtype=K.dtype)
    for i in range(3):
        for l in range(3):
            div_d = np.zeros_like(K)
            cf_d = np.zeros_like(K)
            for j in range(3):
                div_d = div_d + d2B[j, l, i, j]
                cf_d = cf_d + (dBW[l, i, j] * dW[j] + B_over_W[i, j] * d2W[j, l])
            metric_derivative_term = np.einsum('jk...,jk...->...', d_inv[l], dA[i])
            metric_second_term = np.einsum('jk...,jk...->...', inv_gamma, d2A[i, l])
            dM[i, l] = (div_d - 0.5 * metric_derivative_term - 0.5 * metric_second_term
                        - 3.0 * cf_d - 2.0 / 3.0 * d2K[i, l])
    return M, dM


def compute_momentum_constraint(v, params):
    return _momentum_constraint(v, invert_3x3_metric(v.conformal_metric),
                                _Derivatives(params))


def compute_momentum_constraint_and_derivative(v, params):
    return _momentum_constraint_and_derivative(
        v, invert_3x3_metric(v.conformal_metric), _Derivatives(params))


# --------------------------------------------------------------------------
# Evolution equations
# --------------------------------------------------------------------------

def _shift_derivatives(v, D):
    """d_beta[i, j] = d_j beta^i (shift_and_lapse.py:14-40)."""
    return np.swapaxes(D.d1('beta', v.shift), 0, 1)


def _rhs_gamma(v, params, D):
    """spatial_metric.py:17-63."""
    gamma = v.conformal_metric
    gs = _shift_derivatives(v, D)
    first = compute_shift_advection(gamma, v.shift, params)
    second = np.einsum('mi...,mj...->ij...', gamma, gs)
    third = np.einsum('mj...,mi...->ij...', gamma, gs)
    div = np.einsum('ii...->...', gs)
    fourth = -2.0 / 3.0 * gamma * div
    fifth = -2.0 * v.lapse * v.traceless_K
    return first + second + third + fourth + fifth + ko_dissipation(gamma, params)


def _rhs_W(v, params, D):
    """spatial_metric.py:66-106."""
    W = v.conformal_factor
    first = compute_shift_advection(W, v.shift, params)
    second = (1.0 / 3.0) * v.lapse * W * v.trace_K
    div = np.einsum('ii...->...', _shift_derivatives(v, D))
    third = -(1.0 / 3.0) * W * div
    return first + second + third + ko_dissipation(W, params)


def _rhs_K(v, params, inv_gamma, D, hess, matter=None):
    """extrinsic_curvature.py:31-92."""
    alpha, K, A = v.lapse, v.trace_K, v.traceless_K
    first = -np.einsum('ij...,ij...->...', inv_gamma, hess)
    second = alpha * np.einsum('ij...,kl...,ik...,jl...->...', inv_gamma, inv_gamma, A, A)
    third = alpha * K ** 2 / 3.0
    fourth = compute_shift_advection(K, v.shift, params)
    dt_K = first + second + third + fourth
    if matter is not None:
        rho, _, S_ij = matter
        W_floor = np.maximum(v.conformal_factor, W_FLOOR_VALUE)
        stress_trace = np.einsum('ij...,ij...->...', W_floor ** 2 * inv_gamma, S_ij)
        dt_K = dt_K + 4.0 * np.pi * alpha * (rho + stress_trace)
    return dt_K + ko_dissipation(K, params)


def _rhs_A(v, params, inv_gamma, D, hess, ricci, matter=None):
    """extrinsic_curvature.py:119-224."""
    alpha, K, A = v.lapse, v.trace_K, v.traceless_K
    gamma = v.conformal_metric
    c2 = christoffel_symbols_second_kind(inv_gamma, D.d1('gamma', gamma))
    first = alpha * K * A
    second = -2 * alpha * np.einsum('ik...,kl...,lj...->ij...', A, inv_gamma, A)
    third = alpha * ricci - hess
    if matter is not None:
        W_floor = np.maximum(v.conformal_factor, W_FLOOR_VALUE)
        third = third - 8.0 * np.pi * alpha * W_floor ** 2 * matter[2]
    third = traceless_part(third, gamma, inv_gamma)
    gs = _shift_derivatives(v, D)
    fourth = compute_shift_advection(A, v.shift, params)
    fifth = (np.einsum('mi...,mj...->ij...', A, gs) + np.einsum('mj...,mi...->ij...', A, gs))
    div = np.einsum('ii...->...', gs)
    sixth = -2.0 / 3.0 * A * div
    dt_A = first + second + third + fourth + fifth + sixth
    if params.kappa != 0.0:
        M, dM = _momentum_constraint_and_derivative(v, inv_gamma, D)
        if matter is not None:
            S_i = matter[1]
            dS = np.stack([np.stack([diff1_physical(S_i[i], d, params) for d in range(3)])
                           for i in range(3)])   # (i, l) = d_l S_i
            M = M - 8.0 * np.pi * S_i
            dM = dM - 8.0 * np.pi * dS
        DjMi = dM - np.einsum('kij...,k...->ij...', c2, M)
        DiMj = np.swapaxes(DjMi, 0, 1)
        dt_A = dt_A + params.kappa / 2 * alpha * (DjMi + DiMj)
    # kappa == 0: the reference adds kappa/2 * alpha * (...) == +-0 exactly.
    return dt_A + ko_dissipation(A, params)


def _rhs_Gamma(v, params, inv_gamma, D, matter=None):
    """conformal_connection.py:19-178."""
    alpha, A, W = v.lapse, v.traceless_K, v.conformal_factor
    W_floor = np.maximum(W, W_FLOOR_VALUE)
    dW = D.d1('W', W)
    da = D.d1('alpha', alpha)
    dK = D.d1('K', v.trace_K)
    gs = _shift_derivatives(v, D)
    div = np.einsum('ii...->...', gs)
    first = -4 / 3 * alpha * np.einsum('ij...,j...->i...', inv_gamma, dK)
    c2 = christoffel_symbols_second_kind(inv_gamma, D.d1('gamma', v.conformal_metric))
    A_up = np.einsum('ik...,jl...,kl...->ij...', inv_gamma, inv_gamma, A)
    second = 2 * alpha * np.einsum('ijk...,jk...->i...', c2, A_up)
    third = -6 * alpha / W_floor * np.einsum('ij...,j...->i...', A_up, dW)
    fourth = -2 * np.einsum('ij...,j...->i...', A_up, da)
    fifth = compute_shift_advection(v.conformal_connection, v.shift, params)
    sixth = (2.0 / 3.0) * v.conformal_connection * div
    seventh = -np.einsum('m...,im...->i...', v.conformal_connection, gs)
    ddb = D.d2('beta', v.shift)          # (m, n, i): inner m, outer n
    # reference d2_shift[i, m, n] = diff1(d_shift[i, n], m): inner n, outer m
    d2_shift = np.einsum('nmi...->imn...', ddb)
    eighth = np.einsum('mn...,imn...->i...', inv_gamma, d2_shift)
    div_deriv = np.zeros((3,) + v.trace_K.shape, dtype=v.trace_K.dtype)
    for m in range(3):
        for n in range(3):
            # m == n: D2; else diff1(d_shift[n, n], m): inner n, outer m
            div_deriv[m] = div_deriv[m] + ddb[n, m, n]
    ninth = (1.0 / 3.0) * np.einsum('im...,m...->i...', inv_gamma, div_deriv)
    dt = first + second + third + fourth + fifth + sixth + seventh + eighth + ninth
    out = dt + ko_dissipation(v.conformal_connection, params)
    if matter is not None:
        out = out + -16.0 * np.pi * alpha * np.einsum('ij...,j...->i...', inv_gamma, matter[1])
    return out


def _rhs_lapse(v, params):
    """shift_and_lapse.py:65-97."""
    if params.gauge == 0:
        slicing = -np.power(v.lapse, 2) * v.trace_K
    else:
        slicing = -2.0 * v.lapse * v.trace_K
    adv = compute_shift_advection(v.lapse, v.shift, params)
    return slicing + adv + ko_dissipation(v.lapse, params)


def _rhs_shift(v, params):
    """shift_and_lapse.py:100-132 and time_evolve.py:30-38."""
    if params.zero_shift != 0:
        return 0.0 * v.shift
    adv = compute_shift_advection(v.shift, v.shift, params)
    dt_beta = params.g * v.conformal_connection + adv + -params.eta * v.shift
    return dt_beta + ko_dissipation(v.shift, params)


# --------------------------------------------------------------------------
# Sommerfeld faces (evolution/boundaries.py, evolution/coordinates.py)
# --------------------------------------------------------------------------

def grid_coordinates(shape, params):
    """x_min + dx * index per axis, broadcast (coordinates.py:6-28)."""
    mins = (params.x_min, params.y_min, params.z_min)
    out = []
    for d in range(3):
        a = mins[d] + params.dx * np.arange(shape[d], dtype=np.float64)
        out.append(a.reshape((1,) * d + (shape[d],) + (1,) * (2 - d)))
    return tuple(out)


def radial_derivative(field, params):
    """(x^i / r) d_i field (boundaries.py:29-41)."""
    lead = field.ndim - 3
    d = [diff1_physical(field, lead + k, params) for k in range(3)]
    X, Y, Z = grid_coordinates(field.shape[-3:], params)
    r = np.sqrt(X ** 2 + Y ** 2 + Z ** 2)
    r_safe = np.where(r > 0.0, r, 1.0)
    return (X * d[0] + Y * d[1] + Z * d[2]) / r_safe


def sommerfeld_boundary_mask(shape, params):
    """Union of active Sommerfeld face planes (boundaries.py:44-72)."""
    idx = (np.arange(shape[0])[:, None, None], np.arange(shape[1])[None, :, None],
           np.arange(shape[2])[None, None, :])
    start = params.guard_cells
    counts = shape if params.owned_shape is None else params.owned_shape
    codes = (params.xl_bc, params.xr_bc, params.yl_bc, params.yr_bc,
             params.zl_bc, params.zr_bc)
    mask = np.zeros(shape, dtype=bool)
    owned = np.ones(shape, dtype=bool)
    for d in range(3):
        end = start + counts[d]
        lower = codes[2 * d] == SOMMERFELD_BC and params.owned_faces[2 * d]
        upper = codes[2 * d + 1] == SOMMERFELD_BC and params.owned_faces[2 * d + 1]
        if params.owned_shape is not None:
            lower = lower and counts[d] > 1
            upper = upper and counts[d] > 1
        mask = mask | (lower & (idx[d] == start)) | (upper & (idx[d] == end - 1))
        owned = owned & (idx[d] >= start) & (idx[d] < end)
    return mask & owned


def sommerfeld(field, rhs, asymptotic_value, params):
    """Replace the RHS on active faces by -(d_r u + (u - u0)/r) (boundaries.py:75-92)."""
    X, Y, Z = grid_coordinates(field.shape[-3:], params)
    r = np.sqrt(X ** 2 + Y ** 2 + Z ** 2)
    r_safe = np.where(r > 0.0, r, 1.0)
    s_rhs = -(radial_derivative(field, params) + (field - asymptotic_value) / r_safe)
    mask = sommerfeld_boundary_mask(field.shape[-3:], params)
    mask = mask.reshape((1,) * (field.ndim - 3) + mask.shape)
    return np.where(mask, s_rhs, rhs)


def has_sommerfeld(params):
    return any(c == SOMMERFELD_BC for c in (params.xl_bc, params.xr_bc, params.yl_bc,
                                           params.yr_bc, params.zl_bc, params.zr_bc))


def apply_sommerfeld_boundaries(v, rhs, params):
    """boundaries.py:95-172."""
    if not has_sommerfeld(params):
        return rhs
    flat = np.eye(3)[:, :, None, None, None]
    return BSSNVariables(
        conformal_metric=sommerfeld(v.conformal_metric, rhs.conformal_metric, flat, params),
        conformal_factor=sommerfeld(v.conformal_factor, rhs.conformal_factor, 1.0, params),
        traceless_K=sommerfeld(v.traceless_K, rhs.traceless_K, 0.0, params),
        trace_K=sommerfeld(v.trace_K, rhs.trace_K, 0.0, params),
        conformal_connection=sommerfeld(v.conformal_connection, rhs.conformal_connection,
                                        0.0, params),
        lapse=sommerfeld(v.lapse, rhs.lapse, 1.0, params),
        shift=sommerfeld(v.shift, rhs.shift, 0.0, params),
    )


# --------------------------------------------------------------------------
# RHS, projections, RK4 (evolution/time_evolve.py)
# --------------------------------------------------------------------------

def as_float64(v):
    """FP64 fields; complex fields stay complex128, so complex-step
    differentiation (u + i h du) runs through the same code."""
    return BSSNVariables(*(np.asarray(f, dtype=np.result_type(f, np.float64)) for f in v))


def compute_bssn_rhs(v, params):
    """time_evolve.py:41-57."""
    v = as_float64(v)
    D = _Derivatives(params)
    inv_gamma = invert_3x3_metric(v.conformal_metric)
    hess = _w2_lapse_hessian(v, inv_gamma, D)
    ricci = _w2_ricci(v, inv_gamma, D)
    rhs = BSSNVariables(
        conformal_metric=_rhs_gamma(v, params, D),
        conformal_factor=_rhs_W(v, params, D),
        traceless_K=_rhs_A(v, params, inv_gamma, D, hess, ricci),
        trace_K=_rhs_K(v, params, inv_gamma, D, hess),
        conformal_connection=_rhs_Gamma(v, params, inv_gamma, D),
        lapse=_rhs_lapse(v, params),
        shift=_rhs_shift(v, params),
    )
    return apply_sommerfeld_boundaries(v, rhs, params)


def compute_bssn_rhs_with_matter(v, params, energy_density, momentum_density, spatial_stress):
    """time_evolve.py:60-91: sources added before the Sommerfeld faces."""
    v = as_float64(v)
    matter = (np.asarray(energy_density, np.float64), np.asarray(momentum_density, np.float64),
              np.asarray(spatial_stress, np.float64))
    D = _Derivatives(params)
    inv_gamma = invert_3x3_metric(v.conformal_metric)
    hess = _w2_lapse_hessian(v, inv_gamma, D)
    ricci = _w2_ricci(v, inv_gamma, D)
    rhs = BSSNVariables(
        conformal_metric=_rhs_gamma(v, params, D),
        conformal_factor=_rhs_W(v, params, D),
        traceless_K=_rhs_A(v, params, inv_gamma, D, hess, ricci, matter),
        trace_K=_rhs_K(v, params, inv_gamma, D, hess, matter),
        conformal_connection=_rhs_Gamma(v, params, inv_gamma, D, matter),
        lapse=_rhs_lapse(v, params),
        shift=_rhs_shift(v, params),
    )
    return apply_sommerfeld_boundaries(v, rhs, params)


def enforce_unit_determinant_conformal_metric(v):
    """time_evolve.py:94-111."""
    det = determinant_3x3_metric(v.conformal_metric)
    return v._replace(conformal_metric=v.conformal_metric * det ** (-1.0 / 3.0))


def eliminate_trace_A(v):
    """time_evolve.py:114-131."""
    inv = invert_3x3_metric(v.conformal_metric)
    return v._replace(traceless_K=traceless_part(v.traceless_K, v.conformal_metric, inv))


def enforce_algebraic_constraints(v):
    """time_evolve.py:134-141."""
    return eliminate_trace_A(enforce_unit_determinant_conformal_metric(as_float64(v)))


def _axpy(v, c, k):
    return BSSNVariables(*(a + c * b for a, b in zip(v, k)))


def rk4_step(v, params):
    """time_evolve.py:144-217."""
    dt = params.dt
    v = enforce_algebraic_constraints(v)
    k1 = compute_bssn_rhs(v, params)
    k2 = compute_bssn_rhs(enforce_algebraic_constraints(_axpy(v, 0.5 * dt, k1)), params)
    k3 = compute_bssn_rhs(enforce_algebraic_constraints(_axpy(v, 0.5 * dt, k2)), params)
    k4 = compute_bssn_rhs(enforce_algebraic_constraints(_axpy(v, dt, k3)), params)
    out = BSSNVariables(*(u + (dt / 6.0) * (a + 2 * b + 2 * c + d)
                          for u, a, b, c, d in zip(v, k1, k2, k3, k4)))
    return enforce_algebraic_constraints(out)


# --------------------------------------------------------------------------
# Constraints (bssn/constraints.py:262-412, diagnostics/constraints.py)
# --------------------------------------------------------------------------

class ConstraintViolations(NamedTuple):
    hamiltonian: np.ndarray
    momentum: np.ndarray
    det_gamma: np.ndarray
    trace_A: np.ndarray
    gamma_condition: np.ndarray


def compute_hamiltonian_constraint(v, params):
    """constraints.py:271-303."""
    v = as_float64(v)
    inv = invert_3x3_metric(v.conformal_metric)
    ricci = _w2_ricci(v, inv, _Derivatives(params))
    r = trace_tensor(ricci, inv)
    A2 = np.einsum('ik...,jl...,ij...,kl...->...', inv, inv, v.traceless_K, v.traceless_K)
    return r + 2 / 3 * v.trace_K ** 2 - A2


def compute_det_gamma_violation(v):
    return determinant_3x3_metric(np.asarray(v.conformal_metric, np.float64)) - 1.0


def compute_trace_A_violation(v):
    inv = invert_3x3_metric(np.asarray(v.conformal_metric, np.float64))
    return trace_tensor(np.asarray(v.traceless_K, np.float64), inv)


def compute_gamma_constraint(v, params):
    """constraints.py:353-384."""
    v = as_float64(v)
    D = _Derivatives(params)
    inv = invert_3x3_metric(v.conformal_metric)
    c2 = christoffel_symbols_second_kind(inv, D.d1('gamma', v.conformal_metric))
    return v.conformal_connection - np.einsum('jk...,ijk...->i...', inv, c2)


def compute_all_constraints(v, params):
    """constraints.py:387-412."""
    return ConstraintViolations(
        hamiltonian=compute_hamiltonian_constraint(v, params),
        momentum=compute_momentum_constraint(as_float64(v), params),
        det_gamma=compute_det_gamma_violation(v),
        trace_A=compute_trace_A_violation(v),
        gamma_condition=compute_gamma_constraint(v, params),
    )


def compute_all_constraints_with_matter(v, params, energy_density, momentum_density):
    """constraints.py:415-440."""
    c = compute_all_constraints(v, params)
    return c._replace(hamiltonian=c.hamiltonian - 16.0 * np.pi * np.asarray(energy_density),
                      momentum=c.momentum - 8.0 * np.pi * np.asarray(momentum_density))


def compute_constraint_norms(c):
    """diagnostics/constraints.py:9-36."""
    out = {}
    for name, key in (('hamiltonian', 'hamiltonian'), ('momentum', 'momentum'),
                      ('det_gamma', 'det_gamma'), ('trace_A', 'trace_A'),
                      ('gamma', 'gamma_condition')):
        x = getattr(c, key)
        out[name + '_l2'] = float(np.sqrt(np.mean(x ** 2)))
    for name, key in (('hamiltonian', 'hamiltonian'), ('momentum', 'momentum'),
                      ('det_gamma', 'det_gamma'), ('trace_A', 'trace_A'),
                      ('gamma', 'gamma_condition')):
        out[name + '_linf'] = float(np.max(np.abs(getattr(c, key))))
    return out
