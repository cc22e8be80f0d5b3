"""Pins of the FMR oracle (oracle/fmr.py) by what the mathematics fixes:
degree-5 Lagrange prolongation reproduces quintics (the reference's
polynomial ghost-shell check, tolerance 2e-9), coincident child vertices copy
the parent exactly, injection inverts the coincident-vertex map, and a flat
hierarchy stays flat through a full synchronized step."""

import numpy as np
import pytest

from oracle import bssn as OB
from oracle import fmr as OF


def _quintic(t):
    return 0.3 - 0.7 * t + 0.11 * t ** 2 + 0.013 * t ** 3 - 0.002 * t ** 4 + 1.3e-4 * t ** 5


@pytest.mark.parametrize("offset,lo", [(0, (4, 5, 4)), (4, (3, 4, 3))])
def test_prolongation_reproduces_quintics(offset, lo):
    spec = OF.FMRPatchSpec(lo, tuple(l + d for l, d in zip(lo, (4, 3, 3))))
    pshape = (16 + 2 * offset, 15 + 2 * offset, 14 + 2 * offset)
    mats = OF.interpolation_matrices(pshape, spec, offset, periodic=False)
    i, j, k = np.meshgrid(*(np.arange(n, dtype=float) - offset for n in pshape), indexing="ij")
    parent = _quintic(i) * _quintic(0.5 * j + 1) + _quintic(k - 2)
    child = OF.prolongate(parent, mats)
    g = spec.ghost_width
    ti, tj, tk = np.meshgrid(*(spec.coarse_lo[d] + (np.arange(child.shape[d]) - g) / 2 for d in range(3)),
                             indexing="ij")
    exact = _quintic(ti) * _quintic(0.5 * tj + 1) + _quintic(tk - 2)
    assert np.abs(child - exact).max() <= 2e-9 * np.abs(exact).max()


def test_coincident_vertices_copy_and_injection_inverts():
    rng = np.random.default_rng(3)
    spec = OF.FMRPatchSpec((2, 3, 2), (6, 5, 6))
    pshape = (10, 9, 11)
    mats = OF.interpolation_matrices(pshape, spec, 0, periodic=True)
    parent = rng.standard_normal(pshape)
    child = OF.prolongate(parent, mats)
    g = spec.ghost_width
    np.testing.assert_array_equal(child[g::2, g::2, g::2][:5, :3, :5],
                                  parent[2:7, 3:6, 2:7])
    v = OB.BSSNVariables(*(np.broadcast_to(parent, s + pshape).copy() for s in
                           ((3, 3), (), (3, 3), (), (3,), (), (3,))))
    c = OB.BSSNVariables(*(np.broadcast_to(child, s + child.shape).copy() for s in
                           ((3, 3), (), (3, 3), (), (3,), (), (3,))))
    back = OF.restrict_to_coarse(OB.BSSNVariables(*(np.zeros_like(f) for f in v)), c, spec, 0)
    np.testing.assert_array_equal(back.lapse[2:7, 3:6, 2:7], parent[2:7, 3:6, 2:7])
    assert back.lapse[0, 0, 0] == 0.0


def test_flat_hierarchy_stays_flat():
    spec = OF.FMRPatchSpec((2, 2, 2), (5, 5, 5))
    h = OF.FMRHierarchySpec((spec,))
    p = OB.BSSNParameters(dx=0.1, dt=0.025, nu=0.1)

    def flat(shape):
        z = np.zeros(shape)
        return OB.BSSNVariables(np.broadcast_to(np.eye(3)[:, :, None, None, None], (3, 3) + shape).copy(),
                                z + 1, np.zeros((3, 3) + shape), z.copy(), np.zeros((3,) + shape), z + 1,
                                np.zeros((3,) + shape))
    states = (flat((8, 8, 8)), flat(tuple(n + 8 for n in OF.fine_active_shape(spec))))
    params = (p, p._replace(dx=0.05))
    out = OF.fmr_rk4_step(states, params, h)
    for s in out:
        assert np.abs(s.lapse - 1).max() < 1e-14
        assert np.abs(s.traceless_K).max() < 1e-14
