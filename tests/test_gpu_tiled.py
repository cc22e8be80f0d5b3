"""GPU parity of the intra-GPU tiled path (ghost fill + tile-local kernels).

* ``bssn_tiled_ghost_fill`` against the oracle's halo sweep on random fields,
  with endpoints, mixed faces and every exterior rule: copies are exact, the
  linear extrapolation may contract to a fused multiply-add and corners
  extrapolate extrapolated values, so 1e-14 relative to the field's scale;
* tiled RHS with Sommerfeld faces, MAD, kappa and an owned upper endpoint
  against the serial oracle RHS: 1e-10 (the RHS bar);
* three tiled RK4 steps against three serial oracle steps: 1e-12 per field,
  and the constraint norms over owned samples against the serial norms.
"""

import numpy as np
import pytest
import torch

from oracle import bssn as OB
from oracle import tiled as OT
from tests.parity_utils import rel_errors, smooth_state

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def T():
    from jax_bssn_amd import tiled
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    return tiled


def _jp(p):
    import jax_bssn_amd as J
    return J.BSSNParameters(**p._asdict())


def _jl(T, layout):
    return T.TileLayout(tuple(layout.counts), tuple(layout.tile_shape), layout.guard_cells,
                        tuple(layout.upper_endpoints))


def _tile_major(field, layout):
    """Reference-shaped (ncomp, cx, cy, cz, *alloc) -> (ntiles, ncomp, *alloc)."""
    nc = field.shape[0]
    return np.ascontiguousarray(np.swapaxes(field.reshape((nc, -1) + layout.allocated_shape), 0, 1))


@pytest.mark.parametrize("counts,widths,g,ends,bcs,exterior", [
    ((2, 3, 1), (4, 4, 5), 4, (False, False, False), (1, 1, 0, 0, 1, 1), "linear"),
    ((3, 1, 2), (4, 6, 4), 4, (True, False, True), (1, 1, 1, 1, 0, 0), "linear"),
    ((2, 2, 2), (5, 4, 4), 4, (False, True, False), (0, 1, 1, 0, 1, 0), "constant"),
    ((2, 1, 1), (4, 8, 8), 3, (False, False, False), (1, 1, 1, 1, 1, 1), "preserve"),
    ((1, 2, 1), (6, 4, 7), 4, (True, False, False), (0, 0, 0, 0, 0, 0), "linear"),
])
def test_ghost_fill_matches_oracle_sweep(T, counts, widths, g, ends, bcs, exterior):
    layout = OT.TileLayout(counts, widths, g, ends)
    rng = np.random.default_rng(sum(counts) + g)
    field = rng.standard_normal((3,) + tuple(counts) + layout.allocated_shape)
    names = ("xl_bc", "xr_bc", "yl_bc", "yr_bc", "zl_bc", "zr_bc")
    p = OB.BSSNParameters(**dict(zip(names, bcs)))
    want = _tile_major(OT.refresh_tiled_field(field, p, layout, exterior), layout)
    got = torch.tensor(_tile_major(field, layout), device="cuda")
    T.refresh_tiled_field(got, _jp(p), _jl(T, layout), exterior)
    err = np.abs(got.cpu().numpy() - want).max() / np.abs(want).max()
    print("ghost fill", counts, widths, g, ends, bcs, exterior, f"{err:.1e}", flush=True)
    assert err <= 1e-14


def test_tiled_rhs_with_faces_mad_kappa_and_endpoint(T):
    layout = OT.TileLayout((4, 1, 1), (4, 8, 8), 4, (True, False, False))
    params = OB.BSSNParameters(dx=.2, dt=.001, x_min=-2., y_min=-1., z_min=-1., xl_bc=1, xr_bc=1, zl_bc=1,
                               zr_bc=1, mad_q=.35, kappa=.1)
    v = smooth_state((17, 8, 8), seed=3)
    tiled = T.initialize_tiled_bssn(v, _jp(params), _jl(T, layout))
    ev = T.TiledEvolver(tiled, _jp(params), _jl(T, layout))
    rhs = T.gather_owned(ev.rhs(), _jl(T, layout))
    from jax_bssn_amd.variables import unpack
    got = unpack(rhs, np.zeros(1))
    errs = rel_errors(got, OB.compute_bssn_rhs(v, params))
    print("tiled rhs", " ".join(f"{k}={e:.1e}" for k, e in errs.items()), flush=True)
    assert max(errs.values()) <= 1e-10


@pytest.mark.parametrize("counts,widths,params", [
    ((2, 1, 1), (4, 4, 4), OB.BSSNParameters(dx=.2, dt=.001, nu=.02, kappa=.1)),
    ((2, 2, 1), (8, 8, 16), OB.BSSNParameters(dx=.1, dt=.02, nu=.05, eta=1.0, g=0.75, gauge=1,
                                              xl_bc=1, xr_bc=1, zl_bc=1, zr_bc=1, x_min=-.8, z_min=-.8)),
])
def test_tiled_steps_match_serial(T, counts, widths, params):
    layout = OT.TileLayout(counts, widths, 4)
    v = smooth_state(layout.active_shape, seed=4)
    ev = T.TiledEvolver(T.initialize_tiled_bssn(v, _jp(params), _jl(T, layout)), _jp(params), _jl(T, layout))
    ref = v
    from jax_bssn_amd.variables import unpack
    for step in range(3):
        ev.step(1)
        ref = OB.rk4_step(ref, params)
        errs = rel_errors(unpack(ev.global_packed(), np.zeros(1)), ref)
        print("tiled step", counts, step, " ".join(f"{k}={e:.1e}" for k, e in errs.items()), flush=True)
        assert max(errs.values()) <= 1e-12
    norms = ev.constraint_norms()
    want = OB.compute_constraint_norms(OB.compute_all_constraints(ref, params))
    for k, w in want.items():
        assert abs(norms[k] - w) <= 1e-9 * abs(w) + 1e-14, (k, norms[k], w)
    assert norms["det_gamma_linf"] < 3e-14 and norms["trace_A_linf"] < 3e-14


def test_reference_shaped_tiled_step_api(T):
    layout = OT.TileLayout((2, 1, 1), (4, 4, 4), 4)
    params = OB.BSSNParameters(dx=.2, dt=.001, nu=.02)
    v = smooth_state((8, 4, 4), seed=6)
    ref_tiled = OT.refresh_bssn_halos(OB.BSSNVariables(*(OT.tile_from_global(f, layout) for f in v)), params,
                                      layout)
    want = OT.tiled_rk4_step(ref_tiled, params, layout)
    import jax_bssn_amd as J
    got = T.tiled_rk4_step(J.BSSNVariables(*ref_tiled), _jp(params), _jl(T, layout))
    for a, b in zip(got, want):
        a = OT.global_from_tiles(np.asarray(a), layout)
        b = OT.global_from_tiles(np.asarray(b), layout)
        assert np.abs(a - b).max() <= 1e-12 * max(1.0, np.abs(b).max())
