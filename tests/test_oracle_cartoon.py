"""Pins of the axisymmetric Cartoon oracle (oracle/cartoon.py) with the
reference's own checks (tests/Cartoon/test_axisymmetric_cartoon.py:146-417)
and what the mathematics fixes: six-point Lagrange reproduces quintics, the
reconstructed centre plane is the compact state, a rotating vector field
comes back in Cartesian components, the scalar reconstruction converges above
fifth order and the compact RHS converges to the full 3-D RHS of an
axisymmetric state."""

import numpy as np
import pytest

from oracle import bssn as OB
from oracle import cartoon as OC


def _params(nrho, nz, dx, z_min=None):
    if z_min is None:
        z_min = -(nz - 1) * dx / 2.0
    return OB.BSSNParameters(dx=dx, dt=0.05 * dx, nu=0.0, gauge=1, xl_bc=0, xr_bc=1, yl_bc=0, yr_bc=0,
                             zl_bc=0, zr_bc=0, x_min=-3.5 * dx, y_min=-4.0 * dx, z_min=z_min, mad_q=1.0)


def minkowski(nrho=12, nz=17):
    shape = (nrho + OC.GHOSTS, 1, nz)
    s = np.zeros(shape)
    vec = np.zeros((3,) + shape)
    metric = np.eye(3)[:, :, None, None, None] * np.ones((3, 3) + shape)
    return OB.BSSNVariables(metric, np.ones(shape), np.zeros_like(metric), s, vec, np.ones(shape), vec.copy())


def _set_positive(field, values):
    f = np.array(field, copy=True)
    f[..., OC.GHOSTS:, 0, :] = values
    return f


def test_lagrange6_reproduces_quintics_and_flags_outside():
    x = np.arange(11.0)
    f = 0.2 - x + 0.3 * x ** 2 - 0.02 * x ** 3 + 1e-3 * x ** 4 - 2e-5 * x ** 5
    q = np.array([0.0, 0.3, 2.5, 5.75, 9.9, 10.0])
    exact = 0.2 - q + 0.3 * q ** 2 - 0.02 * q ** 3 + 1e-3 * q ** 4 - 2e-5 * q ** 5
    np.testing.assert_allclose(OC.lagrange6(f, q, axis=0), exact, rtol=1e-12, atol=1e-12)
    assert np.isnan(OC.lagrange6(f, np.array([-0.1, 10.5]), axis=0)).all()


def test_reconstructed_centre_plane_equals_filled_compact_state():
    nrho, nz, dx = 16, 17, 0.15
    v = minkowski(nrho, nz)
    rho = (np.arange(nrho) + 0.5) * dx
    z = (np.arange(nz) - (nz - 1) / 2.0) * dx
    R, Z = np.meshgrid(rho, z, indexing="ij")
    shift = np.stack((R * (1 + 0.1 * Z), R * (0.3 - 0.05 * Z), 0.2 + R ** 2))
    m = np.zeros((3, 3, nrho, nz))
    m[0, 0] = 1.0 + 0.02 * R ** 2
    m[1, 1] = 1.0 - 0.01 * R ** 2
    m[2, 2] = 1.0
    m[0, 1] = m[1, 0] = 0.004 * R ** 2
    m[0, 2] = m[2, 0] = 0.02 * R
    m[1, 2] = m[2, 1] = -0.015 * R
    v = OC.fill_axisymmetric_ghosts(v._replace(shift=_set_positive(v.shift, shift),
                                               conformal_metric=_set_positive(v.conformal_metric, m)))
    sup = OC.reconstruct_axisymmetric_support(v, _params(nrho, nz, dx))
    for cf, sf in zip(v, sup):
        np.testing.assert_allclose(sf[..., :, OC.CENTER, :], cf[..., :, 0, :], atol=2e-13, rtol=2e-13)
        assert np.isfinite(sf).all()


def test_azimuthal_vector_rotates_into_cartesian_components():
    nrho, nz, dx = 20, 17, 0.1
    v = minkowski(nrho, nz)
    rho = (np.arange(nrho) + 0.5) * dx
    z = (np.arange(nz) - (nz - 1) / 2.0) * dx
    R, Z = np.meshgrid(rho, z, indexing="ij")
    vec = np.stack((R * (1.0 + 0.1 * Z), 0.3 * R, 0.2 + 0.05 * R ** 2))
    v = OC.fill_axisymmetric_ghosts(v._replace(conformal_connection=_set_positive(v.conformal_connection, vec)))
    sup = OC.reconstruct_axisymmetric_support(v, _params(nrho, nz, dx))
    x = (np.arange(nrho + 4) - 3.5) * dx
    y = (np.arange(9) - 4.0) * dx
    X, Y = np.meshgrid(x, y, indexing="ij")
    r = np.sqrt(X ** 2 + Y ** 2)
    c, s = X / r, Y / r
    zg, rg = z[None, None, :], r[:, :, None]
    ref = np.stack((rg * (1.0 + 0.1 * zg), 0.3 * rg * np.ones_like(zg), (0.2 + 0.05 * rg ** 2) * np.ones_like(zg)))
    ex = c[..., None] * ref[0] - s[..., None] * ref[1]
    ey = s[..., None] * ref[0] + c[..., None] * ref[1]
    inner = (slice(None, -5), slice(None), slice(None))
    np.testing.assert_allclose(sup.conformal_connection[0][inner], ex[inner], atol=3e-12)
    np.testing.assert_allclose(sup.conformal_connection[1][inner], ey[inner], atol=3e-12)


def test_scalar_reconstruction_converges_above_fifth_order():
    errors = []
    for nrho in (24, 32):
        dx = 2.0 / nrho
        nz = 2 * nrho
        p = _params(nrho, nz, dx)
        v = minkowski(nrho, nz)
        rho = (np.arange(nrho) + 0.5) * dx
        z = p.z_min + np.arange(nz) * dx
        R, Z = np.meshgrid(rho, z, indexing="ij")
        v = OC.fill_axisymmetric_ghosts(
            v._replace(conformal_factor=_set_positive(v.conformal_factor, 1.0 + 0.01 * np.exp(-(R ** 2 + Z ** 2)))))
        sup = OC.reconstruct_axisymmetric_support(v, p).conformal_factor
        x = (np.arange(nrho + 4) - 3.5) * dx
        y = (np.arange(9) - 4.0) * dx
        X, Y, Z3 = np.meshgrid(x, y, z, indexing="ij")
        exact = 1.0 + 0.01 * np.exp(-(X ** 2 + Y ** 2 + Z3 ** 2))
        errors.append(np.abs(sup[4:-5, :, 4:-4] - exact[4:-5, :, 4:-4]).max())
    assert np.log(errors[0] / errors[1]) / np.log(32 / 24) > 5.3


def test_minkowski_rhs_and_step_are_stationary():
    v = minkowski(10, 17)
    p = _params(10, 17, 0.2)
    for f in OC.compute_axisymmetric_rhs(v, p):
        np.testing.assert_allclose(f, 0.0, atol=2e-13)
    for a, b in zip(v, OC.axisymmetric_rk4_step(v, p)):
        np.testing.assert_allclose(b, a, atol=2e-13)


def smooth_axisymmetric_state(X, Y, Z):
    """The reference's _smooth_axisymmetric_state (test_axisymmetric_cartoon.py:343-366)."""
    env = np.exp(-(X ** 2 + Y ** 2 + Z ** 2))
    shape = X.shape
    metric = np.eye(3)[:, :, None, None, None] * np.ones((3, 3) + shape)
    radial, az, vert = 0.01 * env, 0.006 * env, 0.004 * env
    shift = np.stack((radial * X - az * Y, radial * Y + az * X, vert * Z))
    return OB.BSSNVariables(metric, 1.0 + 0.01 * env, np.zeros_like(metric), 0.005 * env, 0.5 * shift,
                            1.0 - 0.02 * env, shift)


def full_and_compact(full_nx):
    dx = 4.0 / full_nx
    x = (np.arange(full_nx) - (full_nx - 1) / 2.0) * dx
    y = (np.arange(full_nx + 1) - full_nx / 2.0) * dx
    z = x
    X0, Z0 = np.meshgrid(x, z, indexing="ij")
    plane = smooth_axisymmetric_state(X0[:, None, :], np.zeros_like(X0[:, None, :]), Z0[:, None, :])
    compact = OC.compact_axisymmetric_state(plane)
    cp = _params(full_nx // 2, full_nx, dx, float(z[0]))._replace(dt=0.01 * dx)
    return x, y, z, dx, compact, cp


def test_compact_rhs_converges_to_full_3d_rhs():
    errors = []
    for full_nx in (24, 32):
        x, y, z, dx, compact, cp = full_and_compact(full_nx)
        X, Y, Z = np.meshgrid(x, y, z, indexing="ij")
        full = smooth_axisymmetric_state(X, Y, Z)
        fp = OB.BSSNParameters(dx=dx, dt=0.01 * dx, nu=0.0, gauge=1, x_min=float(x[0]), y_min=float(y[0]),
                               z_min=float(z[0]))
        full_rhs = OB.compute_bssn_rhs(full, fp)
        comp_rhs = OC.compute_axisymmetric_rhs(compact, cp)
        errs = []
        for ff, cf in zip(full_rhs, comp_rhs):
            ref = ff[(..., slice(full_nx // 2, -4), full_nx // 2, slice(4, -4))]
            cand = cf[(..., slice(OC.GHOSTS, -4), 0, slice(4, -4))]
            errs.append(np.abs(ref - cand).max())
        errors.append(errs)
    assert np.log(max(errors[0]) / max(errors[1])) / np.log(32 / 24) > 3.4
