"""Pins of the tiled oracle (oracle/tiled.py) by the reference's own checks
(tests/test_parallel.py:140-174): a tiled RHS with Sommerfeld faces, MAD, the
kappa term and an owned upper endpoint equals the serial RHS, and three tiled
RK4 steps equal three serial steps, at the reference's 3e-11, with the
algebraic constraints held to 3e-14 on owned samples."""

import numpy as np

from oracle import bssn as OB
from oracle import tiled as OT
from tests.parity_utils import smooth_state


def _tile_state(v, layout, params):
    return OT.refresh_bssn_halos(OB.BSSNVariables(*(OT.tile_from_global(f, layout) for f in v)), params, layout)


def _close(a, b, tol=3e-11):
    a, b = np.asarray(a), np.asarray(b)
    assert np.abs(a - b).max() <= tol * max(1.0, np.abs(b).max())


def test_tiled_rhs_with_faces_mad_kappa_and_endpoint_equals_serial():
    layout = OT.TileLayout((4, 1, 1), (4, 8, 8), 4, (True, False, False))
    params = OB.BSSNParameters(dx=.2, dt=.001, x_min=-2., y_min=-1., z_min=-1., xl_bc=1, xr_bc=1, zl_bc=1,
                               zr_bc=1, mad_q=.35, kappa=.1)
    v = smooth_state((17, 8, 8), seed=3)
    tiled = _tile_state(v, layout, params)
    got = OT.map_tiled(OB.compute_bssn_rhs, tiled, params, layout)
    want = OB.compute_bssn_rhs(v, params)
    for a, b in zip(got, want):
        _close(OT.global_from_tiles(a, layout), b)


def test_three_tiled_steps_equal_serial_steps():
    layout = OT.TileLayout((2, 1, 1), (4, 4, 4), 4)
    params = OB.BSSNParameters(dx=.2, dt=.001, nu=.02, kappa=.1)
    v = smooth_state((8, 4, 4), seed=4)
    tiled = _tile_state(v, layout, params)
    for _ in range(3):
        tiled = OT.tiled_rk4_step(tiled, params, layout)
        v = OB.rk4_step(v, params)
    for a, b in zip(tiled, v):
        _close(OT.global_from_tiles(a, layout), b)
    g = OT.global_from_tiles(tiled.conformal_metric, layout)
    det = np.linalg.det(np.moveaxis(g, (0, 1), (-2, -1)))
    assert np.abs(det - 1).max() < 3e-14


def test_sweep_fills_corners_and_exteriors():
    """Guards of a linear field are exact under linear extrapolation, and
    the periodic exchange reproduces the wrapped global field."""
    layout = OT.TileLayout((2, 3, 1), (4, 4, 5), 3)
    shape = layout.active_shape
    i, j, k = np.meshgrid(*(np.arange(n, dtype=float) for n in shape), indexing="ij")
    lin = 1.0 + 0.5 * i - 0.25 * j + 0.125 * k
    p_somm = OB.BSSNParameters(xl_bc=1, xr_bc=1, yl_bc=1, yr_bc=1, zl_bc=1, zr_bc=1)
    t = OT.refresh_tiled_field(OT.tile_from_global(lin, layout), p_somm, layout)
    g = layout.guard_cells
    for (a, b, c) in [(0, 0, 0), (1, 2, 0)]:
        ii, jj, kk = np.meshgrid(*(np.arange(-g, w + g) + t_ * w for t_, w in zip((a, b, c), layout.tile_shape)),
                                 indexing="ij")
        np.testing.assert_allclose(t[a, b, c], 1.0 + 0.5 * ii - 0.25 * jj + 0.125 * kk, atol=1e-13)
    rng = np.random.default_rng(1)
    f = rng.standard_normal(shape)
    t = OT.refresh_tiled_field(OT.tile_from_global(f, layout), OB.BSSNParameters(), layout)
    ii, jj, kk = np.meshgrid(*(np.arange(-g, w + g) + w for w in layout.tile_shape), indexing="ij")
    np.testing.assert_array_equal(t[1, 1, 0], f[ii % shape[0], jj % shape[1], kk % shape[2]])
