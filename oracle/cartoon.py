"""FP64 NumPy restatement of the reference's axisymmetric Cartoon path.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

Follows reference JAX_BSSN/cartoon/:
  * parity tables and reflection (symmetry.py:14-71), vector parity
    (-1, -1, +1) for a rotation by pi about z (axisymmetry/reconstruction.py:31);
  * six-point non-periodic Lagrange interpolation (interpolation.py:9-87);
  * compact storage: four parity ghosts then positive radial half-cell
    centres, singleton y (reconstruction.py:37-145);
  * the 1/r outer buffer (:148-170), support geometry (:173-182), radial
    interpolation (:185-209), rotation by phi about z (:212-252),
    reconstruction on nine y planes (:255-292), projection of the centre
    plane (:295-304);
  * the stage-wise RK4 step (axisymmetry/evolution.py:18-65).
"""

import numpy as np

from . import bssn as OB

GHOSTS = 4
SUPPORT = 9
CENTER = 4
OUTER = 3
VECTOR_PARITY = (-1.0, -1.0, 1.0)


def _parity(field):
    rank = field.ndim - 3
    v = np.asarray(VECTOR_PARITY)
    if rank == 0:
        return np.asarray(1.0)
    if rank == 1:
        return v
    return v[:, None] * v[None, :]


def reflect(positive, parity, axis, count=None):
    """symmetry.py:55-71: mirrored first ``count`` samples times parity, then
    the positive half."""
    axis = axis % positive.ndim
    parity = np.asarray(parity, dtype=float)
    parity = parity.reshape(parity.shape + (1,) * (positive.ndim - parity.ndim))
    idx = [slice(None)] * positive.ndim
    idx[axis] = slice(None, count)
    mirrored = np.flip(positive[tuple(idx)], axis=axis) * parity
    return np.concatenate((mirrored, positive), axis=axis)


def lagrange6(array, q, axis):
    """interpolation.py:9-87: six nodes starting at clip(floor(q) - 2, 0,
    n - 6), NaN outside [0, n - 1]; targets q of any shape replace ``axis``."""
    axis = axis % array.ndim
    n = array.shape[axis]
    q = np.asarray(q, dtype=float)
    base = np.floor(q).astype(int)
    start = np.clip(base - 2, 0, n - 6)
    nodes = start[..., None] + np.arange(6)
    w = np.ones(nodes.shape)
    for k in range(6):
        for m in range(6):
            if k != m:
                w[..., k] *= (q - nodes[..., m]) / (nodes[..., k] - nodes[..., m])
    moved = np.moveaxis(array, axis, -1)
    gathered = np.take(moved, nodes, axis=-1)        # (..., *q.shape, 6)
    out = np.sum(gathered * w, axis=-1)              # (..., *q.shape)
    valid = (q >= 0) & (q <= n - 1)
    out = np.where(valid, out, np.nan)
    lead = array.ndim - 1
    # move target axes back to where ``axis`` was
    perm = list(range(axis)) + list(range(lead, lead + q.ndim)) + list(range(axis, lead))
    return np.transpose(out, perm)


def _positive_plane(field):
    return np.asarray(field)[..., GHOSTS:, 0, :]


def _compact(positive, parity):
    return np.expand_dims(reflect(positive, parity, axis=-2, count=GHOSTS), axis=-2)


def fill_axisymmetric_ghosts(v):
    """reconstruction.py:131-145."""
    return OB.BSSNVariables(*(_compact(_positive_plane(f), _parity(np.asarray(f))) for f in v))


def compact_axisymmetric_state(v):
    """reconstruction.py:116-128: signed x plane (even length, y = 1) ->
    compact storage."""
    nx = np.asarray(v.lapse).shape[0]
    return OB.BSSNVariables(*(_compact(np.asarray(f)[..., nx // 2:, 0, :], _parity(np.asarray(f))) for f in v))


def _outer_buffer(reference, asymptotic, dx, z_min):
    nx, nz = reference.shape[-2:]
    nr = nx - GHOSTS
    rho_edge = (nr - 0.5) * dx
    z = z_min + dx * np.arange(nz)
    rho_buf = rho_edge + dx * np.arange(1, OUTER + 1)
    r_edge = np.sqrt(rho_edge ** 2 + z ** 2)
    r_buf = np.sqrt(rho_buf[:, None] ** 2 + z[None, :] ** 2)
    ratio = r_edge[None, :] / r_buf
    asym = np.asarray(asymptotic, dtype=float)
    asym = asym.reshape(asym.shape + (1, 1))
    buf = asym + (reference[..., -1:, :] - asym) * ratio
    return np.concatenate((reference, buf), axis=-2)


def support_geometry(nx, dx):
    x = (np.arange(nx) - 3.5) * dx
    y = (np.arange(SUPPORT) - 4.0) * dx
    X, Y = np.broadcast_arrays(x[:, None], y[None, :])
    return np.sqrt(X ** 2 + Y ** 2), X, Y


def radial_interpolate(positive, rho, dx, z_min, parity, asymptotic):
    plane = reflect(positive, parity, axis=-2, count=GHOSTS)
    source = _outer_buffer(plane, asymptotic, dx, z_min)
    q = rho / dx + GHOSTS - 0.5
    return lagrange6(source, q, axis=-2)


def transformation_matrix(x, y):
    """reconstruction.py:212-244: Lambda^i_j, columns radial / azimuthal / axial."""
    x, y = np.broadcast_arrays(np.asarray(x, float), np.asarray(y, float))
    rho = np.sqrt(x ** 2 + y ** 2)
    on = rho <= 0
    safe = np.where(on, 1.0, rho)
    c = np.where(on, 1.0, x / safe)
    s = np.where(on, 0.0, y / safe)
    z0, o1 = np.zeros_like(x), np.ones_like(x)
    return np.stack((np.stack((c, s, z0)), np.stack((-s, c, z0)), np.stack((z0, z0, o1))), axis=1)


def reconstruct_axisymmetric_support(v, params):
    nx = np.asarray(v.lapse).shape[0]
    rho, X, Y = support_geometry(nx, params.dx)
    R = transformation_matrix(X[..., None], Y[..., None])
    eye = np.eye(3)

    def prof(f, asym):
        return radial_interpolate(_positive_plane(f), rho, params.dx, params.z_min, _parity(np.asarray(f)), asym)

    rot_v = lambda a: np.einsum("ij...,j...->i...", R, a)
    rot_t = lambda a: np.einsum("ia...,jb...,ab...->ij...", R, R, a)
    return OB.BSSNVariables(
        conformal_metric=rot_t(prof(v.conformal_metric, eye)),
        conformal_factor=prof(v.conformal_factor, 1.0),
        traceless_K=rot_t(prof(v.traceless_K, np.zeros((3, 3)))),
        trace_K=prof(v.trace_K, 0.0),
        conformal_connection=rot_v(prof(v.conformal_connection, np.zeros(3))),
        lapse=prof(v.lapse, 1.0),
        shift=rot_v(prof(v.shift, np.zeros(3))))


def project_axisymmetric_rhs(rhs):
    """reconstruction.py:295-304."""
    return OB.BSSNVariables(*(_compact(np.asarray(f)[..., GHOSTS:, CENTER, :], _parity(np.asarray(f))) for f in rhs))


def compute_axisymmetric_rhs(v, params):
    """axisymmetry/evolution.py:18-28 (full nine-plane RHS, centre kept)."""
    return project_axisymmetric_rhs(OB.compute_bssn_rhs(reconstruct_axisymmetric_support(v, params), params))


def _prepare(v):
    return fill_axisymmetric_ghosts(OB.enforce_algebraic_constraints(v))


def _add(v, k, s):
    return OB.BSSNVariables(*(np.asarray(a) + s * np.asarray(b) for a, b in zip(v, k)))


def axisymmetric_rk4_step(v, params):
    """axisymmetry/evolution.py:36-65."""
    dt = params.dt
    v = _prepare(v)
    k1 = compute_axisymmetric_rhs(v, params)
    k2 = compute_axisymmetric_rhs(_prepare(_add(v, k1, 0.5 * dt)), params)
    k3 = compute_axisymmetric_rhs(_prepare(_add(v, k2, 0.5 * dt)), params)
    k4 = compute_axisymmetric_rhs(_prepare(_add(v, k3, dt)), params)
    new = OB.BSSNVariables(*(np.asarray(f) + (dt / 6.0) * (np.asarray(a) + 2.0 * np.asarray(b)
                                                          + 2.0 * np.asarray(c) + np.asarray(d))
                             for f, a, b, c, d in zip(v, k1, k2, k3, k4)))
    return _prepare(new)
