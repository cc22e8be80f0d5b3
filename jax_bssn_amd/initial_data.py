"""Closed-form initial data of the reference's test and demo configurations
(host NumPy; the evolution takes them as ordinary states).

* ``periodic_coordinate_arrays``  reference tests/initial_data.py:62-73
* ``gauge_wave_analytic_state``   tests/initial_data.py:147-205
* ``linear_wave_initial_data``    demos/gravitational_waves/linear_wave.py:36-56
* ``brill_lindquist_conformal_factor``  utilities/bowen_york_solver.py:52-63
* ``puncture_state``  zero-momentum punctures, W = psi^-2, alpha = W
  (demos/axisymmetric_blackhole_collision/axisymmetric_blackhole_collision.py:118-128)

Every metric here is diagonal, so the inverse and the trace-free split are
written out per component.
"""

import numpy as np

from .variables import BSSNVariables


def periodic_coordinate_arrays(grid_size, domain_length=1.0):
    axis = np.linspace(-domain_length / 2.0, domain_length / 2.0, grid_size, endpoint=False)
    return np.meshgrid(axis, axis, axis, indexing="ij")


def _diagonal_state(gxx, gyy, gzz, W, Kxx, Kyy, Kzz, lapse, connection):
    """BSSN variables from a diagonal conformal metric gamma~ and a diagonal
    curvature K_ij, split as in the reference's test data:
    tr = gamma~^ij K_ij, A = W^2 (K_ij - gamma~_ij tr / 3), K = W^2 tr."""
    shape = np.shape(W)
    g = np.zeros((3, 3) + shape)
    g[0, 0], g[1, 1], g[2, 2] = gxx, gyy, gzz
    trK = W ** 2 * (Kxx / gxx + Kyy / gyy + Kzz / gzz)
    A = np.zeros_like(g)
    for i, (gi, Ki) in enumerate(((gxx, Kxx), (gyy, Kyy), (gzz, Kzz))):
        A[i, i] = W ** 2 * Ki - gi * trK / 3.0
    return BSSNVariables(g, np.asarray(W, dtype=float), A, trK, connection, np.asarray(lapse, dtype=float),
                         np.zeros((3,) + shape))


def gauge_wave_analytic_state(X, Y, Z, t=0.0, amplitude=0.1, wavelength=1.0):
    """Exact harmonic gauge wave H = 1 - A sin(2 pi (x - t) / d)."""
    X = np.asarray(X, dtype=float)
    phase = 2.0 * np.pi * (X - t) / wavelength
    H = 1.0 - amplitude * np.sin(phase)
    W = H ** (-1.0 / 6.0)
    Kxx = -(np.pi * amplitude / wavelength) * np.cos(phase) / np.sqrt(H)
    zero = np.zeros_like(H)
    dH = -(2.0 * np.pi * amplitude / wavelength) * np.cos(phase)
    connection = np.stack(((2.0 / 3.0) * H ** (-5.0 / 3.0) * dH, zero, zero))
    return _diagonal_state(H * W ** 2, W ** 2 + zero, W ** 2 + zero, W, Kxx, zero, zero, np.sqrt(H), connection)


def linear_wave_initial_data(X, Y, Z, amplitude=1.0e-8, wavelength=1.0):
    """Plus-polarised wave h = A sin(k x) with its analytic connection."""
    X = np.asarray(X, dtype=float)
    k = 2.0 * np.pi / wavelength
    h = amplitude * np.sin(k * X)
    dh = k * amplitude * np.cos(k * X)
    W = (1.0 - h * h) ** (-1.0 / 6.0)
    zero = np.zeros_like(h)
    # K~_ij given on the conformal metric in the demo: K_yy = dh/2, K_zz = -dh/2
    gxx, gyy, gzz = W * W, (1.0 + h) * W * W, (1.0 - h) * W * W
    trK = dh / 2.0 / gyy - dh / 2.0 / gzz
    A = np.zeros((3, 3) + h.shape)
    A[1, 1] = W * W * (dh / 2.0 - gyy * trK / 3.0)
    A[2, 2] = W * W * (-dh / 2.0 - gzz * trK / 3.0)
    A[0, 0] = W * W * (-gxx * trK / 3.0)
    g = np.zeros((3, 3) + h.shape)
    g[0, 0], g[1, 1], g[2, 2] = gxx, gyy, gzz
    connection = np.stack((2.0 * h * dh / 3.0 * (1.0 - h * h) ** (-2.0 / 3.0), zero, zero))
    return BSSNVariables(g, W, A, W * W * trK, connection, np.ones_like(h), np.zeros((3,) + h.shape))


def brill_lindquist_conformal_factor(masses, positions, X, Y, Z):
    """psi = 1 + sum_a m_a / (2 |x - x_a|)."""
    psi = np.ones_like(np.asarray(X, dtype=float))
    for m, p in zip(masses, positions):
        r = np.sqrt((X - p[0]) ** 2 + (Y - p[1]) ** 2 + (Z - p[2]) ** 2)
        psi = psi + m / (2.0 * r)
    return psi


def puncture_state(X, Y, Z, masses=(1.0,), positions=((0.0, 0.0, 0.0),), lapse_perturbation=None):
    """Zero-momentum punctures: W = psi^-2, alpha = W, gamma~ = delta,
    A = K = Gamma = beta = 0; ``lapse_perturbation`` = (eps, centre, width)
    adds eps exp(-|x - c|^2 / width^2) to the lapse."""
    W = brill_lindquist_conformal_factor(masses, positions, X, Y, Z) ** -2
    shape = W.shape
    g = np.zeros((3, 3) + shape)
    for i in range(3):
        g[i, i] = 1.0
    lapse = W.copy()
    if lapse_perturbation is not None:
        eps, c, width = lapse_perturbation
        lapse = lapse + eps * np.exp(-((X - c[0]) ** 2 + (Y - c[1]) ** 2 + (Z - c[2]) ** 2) / width ** 2)
    z3 = np.zeros((3,) + shape)
    return BSSNVariables(g, W, np.zeros((3, 3) + shape), np.zeros(shape), z3.copy(), lapse, z3.copy())
