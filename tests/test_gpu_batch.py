"""GPU checks of the batched evolution (shared launches per stage).

* every instance of a batch of uneven periodic gauge waves equals its own
  single-grid evolution after three steps (the same per-point code: FP64
  agrees to 1e-14), and one instance matches the FP64 oracle step by step;
* all tile-shape variants give the same FP64 states;
* a HIP-graph-captured step, replayed three times with eager work and
  synchronisations in between, equals the eager batched step every time;
* a batch mixing periodic gauge waves and Sommerfeld instances (lean and
  face tiles of every tile shape) equals the single-grid evolutions and the
  oracle;
* configurations outside the batched path raise NotImplementedError.
"""

import numpy as np
import pytest
import torch

from oracle import bssn as OB
from oracle import initial_data as OI
from tests.parity_utils import rel_errors

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def J():
    import jax_bssn_amd as J
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    return J


def gauge_batch(J, sizes, seed=0, nu=0.1):
    rng = np.random.default_rng(seed)
    states, params, ref = [], [], []
    for n in sizes:
        amp, d = rng.uniform(0.01, 0.1), rng.uniform(0.5, 1.0)
        X, Y, Z = OI.periodic_coordinate_arrays(n, d)
        v = OI.gauge_wave_analytic_state(X, Y, Z, amplitude=amp, wavelength=d)
        p = OB.BSSNParameters(eta=0.0, kappa=0.0, g=0.0, nu=nu, dx=d / n, dt=0.25 * d / n, zero_shift=1, gauge=0)
        ref.append((v, p))
        from jax_bssn_amd.variables import pack
        states.append(pack(v))
        params.append(J.BSSNParameters(**p._asdict()))
    return states, params, ref


def test_batch_matches_single_grid_and_oracle(J):
    from jax_bssn_amd.batch import BatchEvolver
    states, params, ref = gauge_batch(J, (16, 24, 32, 40, 48))
    b = BatchEvolver(states, params, tile_variant=0, grid_blocks=512)
    singles = [J.Evolver(s.clone(), p) for s, p in zip(states, params)]
    v0, p0 = ref[2]
    for step in range(3):
        b.step(1)
        for e in singles:
            e.step(1)
        v0 = OB.rk4_step(v0, p0)
        errs = rel_errors(b.variables(np.zeros(1))[2], v0)
        assert max(errs.values()) <= 1e-12, (step, errs)
    for i, (bs, e) in enumerate(zip(b.states, singles)):
        d = float((bs - e.state).abs().max() / e.state.abs().max())
        print("batch vs single", i, f"{d:.1e}", flush=True)
        assert d <= 1e-14


def test_tile_variants_agree(J):
    from jax_bssn_amd.batch import BatchEvolver
    states, params, _ = gauge_batch(J, (24, 32, 20), seed=3)
    out = []
    for v in (0, 1, 2):
        b = BatchEvolver([s.clone() for s in states], params, tile_variant=v, grid_blocks=256)
        b.step(2)
        out.append(torch.cat([s.reshape(-1) for s in b.states]))
    for o in out[1:]:
        assert float((o - out[0]).abs().max()) <= 1e-14 * float(out[0].abs().max())


def test_graph_replays_equal_eager_steps(J):
    from jax_bssn_amd.batch import BatchEvolver
    states, params, _ = gauge_batch(J, (32, 16, 24), seed=5)
    eager = BatchEvolver([s.clone() for s in states], params, tile_variant=1, grid_blocks=512)
    graphed = BatchEvolver([s.clone() for s in states], params, tile_variant=1, grid_blocks=512)
    graphed.capture()      # the warm-up step advances the graphed batch by one step
    eager.step(1)
    for rep in range(3):
        graphed.step(1)
        torch.cuda.synchronize()
        _ = torch.ones(1000, device="cuda").sum().item()  # eager work between replays
        eager.step(1)
        torch.cuda.synchronize()
        for a, e in zip(graphed.states, eager.states):
            assert torch.equal(a, e), rep


def test_autotuner_picks_a_valid_variant(J):
    from jax_bssn_amd.batch import TILE_VARIANTS, BatchEvolver
    states, params, _ = gauge_batch(J, (16, 24), seed=7)
    b = BatchEvolver(states, params)
    assert b.tile_variant in TILE_VARIANTS and b.grid_blocks >= 1
    print("autotuned", b.tuning, flush=True)


def test_unsupported_configurations_raise(J):
    from jax_bssn_amd.batch import BatchEvolver
    states, params, _ = gauge_batch(J, (16,), seed=1)
    # 20 points along y with Sommerfeld y faces: not on the 8-point tile rows
    states20, params20, _ = gauge_batch(J, (20,), seed=1)
    with pytest.raises(NotImplementedError, match="tile grid"):
        BatchEvolver(states20, [params20[0]._replace(yl_bc=1, yr_bc=1)], tile_variant=0, grid_blocks=64)
    with pytest.raises(NotImplementedError, match="kappa"):
        BatchEvolver(states, [params[0]._replace(kappa=0.1)], tile_variant=0, grid_blocks=64)
    with pytest.raises(NotImplementedError, match="mad_q"):
        BatchEvolver(states, [params[0]._replace(mad_q=0.5)], tile_variant=0, grid_blocks=64)


def sommerfeld_instances(J, shapes, seed=50):
    from jax_bssn_amd.variables import pack
    from tests.parity_utils import smooth_state
    states, params, ref = [], [], []
    for k, sh in enumerate(shapes):
        v = smooth_state(sh, seed=seed + k)
        p = OB.BSSNParameters(eta=2.0, g=0.75, nu=0.1, dx=0.1, dt=0.025, zero_shift=0, gauge=1, xl_bc=1, xr_bc=1,
                              yl_bc=1, yr_bc=1, zl_bc=1, zr_bc=1, x_min=-0.75, y_min=-0.8, z_min=-0.83)
        states.append(pack(v))
        params.append(J.BSSNParameters(**p._asdict()))
        ref.append((v, p))
    return states, params, ref


@pytest.mark.parametrize("variant", [0, 1, 2])
def test_batch_with_sommerfeld_instances(J, variant):
    """Periodic gauge waves and Sommerfeld instances (faces on every tile
    shape's grid) in one batch: lean tiles and face tiles in their own
    launches; each instance equals its single-grid evolution
    (1e-14) and a Sommerfeld one matches the oracle step by step (1e-12)."""
    from jax_bssn_amd.batch import BatchEvolver
    gs, gp, _ = gauge_batch(J, (16, 24), seed=3)
    ss, sp, sref = sommerfeld_instances(J, ((24, 24, 24), (16, 16, 32), (32, 24, 16)))
    states, params = gs + ss, gp + sp
    print(f"[batch sommerfeld] variant={variant}", flush=True)
    b = BatchEvolver([s.clone() for s in states], params, tile_variant=variant, grid_blocks=256)
    assert b.plan.face_tiles > 0
    singles = [J.Evolver(s.clone(), p) for s, p in zip(states, params)]
    v0, p0 = sref[0]
    for step in range(3):
        b.step(1)
        for e in singles:
            e.step(1)
        v0 = OB.rk4_step(v0, p0)
        errs = rel_errors(b.variables(np.zeros(1))[2], v0)
        assert max(errs.values()) <= 1e-12, (step, errs)
    for i, (bs, e) in enumerate(zip(b.states, singles)):
        d = float((bs - e.state).abs().max() / e.state.abs().max())
        print("batch (sommerfeld) vs single", i, f"{d:.1e}", flush=True)
        assert d <= 1e-14


def test_f32rhs_batch_matches_single_grid_f32rhs(J):
    """FP32-RHS batched steps equal the FP32-RHS single-grid steps of each
    instance (the same per-point code; 1e-12 allows for contraction
    differences between the two kernels), and stay within the FP32-RHS error
    of the FP64 evolution."""
    from jax_bssn_amd.batch import BatchEvolver
    states, params, _ = gauge_batch(J, (16, 24, 32), seed=11)
    b = BatchEvolver([s.clone() for s in states], params, precision="f32rhs", tile_variant=0, grid_blocks=256)
    b.step(2)
    for s, p, bs in zip(states, params, b.states):
        e32 = J.Evolver(s.clone(), p, precision="f32rhs").step(2)
        e64 = J.Evolver(s.clone(), p).step(2)
        scale = float(e64.state.abs().max())
        d32 = float((bs - e32.state).abs().max()) / scale
        d64 = float((bs - e64.state).abs().max()) / scale
        print("f32rhs batch vs single f32rhs", f"{d32:.1e}", "vs f64", f"{d64:.1e}", flush=True)
        assert d32 <= 1e-12
        assert d64 <= 1e-5


def test_sommerfeld_batch_graph_and_f32rhs(J):
    """A batch with Sommerfeld instances (five launches per stage): graph
    replays, with eager work in between, equal the eager batched steps; and
    its FP32-RHS steps agree with each instance's single-grid FP32-RHS steps
    (the same tile shape, so the same FP32 offsets; the face tiles of the two
    paths are separate instantiations of the same code, which in FP64 agree
    to 6e-15 and in FP32 to FP32 rounding: 1e-6) and stay within the
    FP32-RHS error of FP64."""
    from jax_bssn_amd.batch import BatchEvolver
    gs, gp, _ = gauge_batch(J, (16,), seed=9)
    ss, sp, _ = sommerfeld_instances(J, ((24, 24, 24), (16, 16, 32)), seed=60)
    states, params = gs + ss, gp + sp
    print("[batch sommerfeld graph / f32rhs]", flush=True)
    eager = BatchEvolver([s.clone() for s in states], params, tile_variant=0, grid_blocks=256)
    graphed = BatchEvolver([s.clone() for s in states], params, tile_variant=0, grid_blocks=256)
    graphed.capture()
    eager.step(1)
    for rep in range(3):
        graphed.step(1)
        torch.cuda.synchronize()
        _ = torch.ones(1000, device="cuda").sum().item()
        eager.step(1)
        torch.cuda.synchronize()
        for a, e in zip(graphed.states, eager.states):
            assert torch.equal(a, e), rep
    b = BatchEvolver([s.clone() for s in states], params, precision="f32rhs", tile_variant=0, grid_blocks=256)
    b.step(2)
    for s, p, bs in zip(states, params, b.states):
        e32 = J.Evolver(s.clone(), p, precision="f32rhs").step(2)
        e64 = J.Evolver(s.clone(), p).step(2)
        scale = float(e64.state.abs().max())
        d32 = float((bs - e32.state).abs().max()) / scale
        d64 = float((bs - e64.state).abs().max()) / scale
        print("sommerfeld f32rhs batch vs single f32rhs", f"{d32:.1e}", "vs f64", f"{d64:.1e}", flush=True)
        assert d32 <= 1e-6
        assert d64 <= 1e-5
