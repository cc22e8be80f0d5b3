"""GPU checks of the forward-mode differentiation (dual-number kernels).

* the value part of the tangent kernels equals the ordinary kernels
  (1e-10 RHS, the RHS bar; 1e-12 per step), on the three kernel paths:
  periodic tile kernel, tile kernels with Sommerfeld face tiles, and the
  point kernel (Sommerfeld faces off the tile grid);
* a run carrying two directions equals the two single-direction runs
  (1e-13: the same arithmetic per direction);
* the tangent RHS and the tangent evolution agree with central finite
  differences of the ordinary kernels (step h = 1e-4 relative; the
  truncation error is O(h^2), so agreement to 1e-6 relative per field);
* the gradient of the L2 Hamiltonian constraint after a few steps with
  respect to the puncture mass M and a Gaussian lapse amplitude eps agrees
  with central differences of the forward evolution to 1e-6 relative, and
  torch autograd returns it through HamiltonianL2AfterSteps.
Agreement with the reference's jax.grad is parity unpinned (JAX is absent).
"""

import numpy as np
import pytest
import torch

from oracle import bssn as OB
from tests.parity_utils import smooth_state

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def J():
    import jax_bssn_amd as J
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    return J


def _field_rel(a, b):
    groups = ((0, 6), (6, 7), (7, 13), (13, 14), (14, 17), (17, 18), (18, 21))
    out = []
    for lo, hi in groups:
        scale = float(b[lo:hi].abs().max())
        if scale > 1e-14:
            out.append(float((a[lo:hi] - b[lo:hi]).abs().max()) / scale)
    return max(out)


def _cases(J):
    from jax_bssn_amd.variables import pack
    v = smooth_state((16, 16, 16), seed=5)
    p = J.BSSNParameters(eta=0.3, g=0.75, nu=0.05, dx=1 / 16, dt=0.2 / 16, zero_shift=0, gauge=1)
    v2 = smooth_state((16, 17, 18), seed=7)
    p2 = J.BSSNParameters(eta=2.0, g=0.75, nu=0.1, dx=0.1, dt=0.025, zero_shift=0, gauge=1, xl_bc=1, xr_bc=1,
                          yl_bc=1, yr_bc=1, zl_bc=1, zr_bc=1, x_min=-0.8, y_min=-0.8, z_min=-0.9)
    v3 = smooth_state((16, 16, 24), seed=9)
    p3 = p2._replace(dx=0.1, dt=0.025, z_min=-1.2)
    out = []
    for vv, pp, shape, seed in ((v, p, (16, 16, 16), 15), (v2, p2, (16, 17, 18), 17), (v3, p3, (16, 16, 24), 19)):
        u = pack(vv)
        # a smooth direction (the difference of two smooth states), so that
        # the finite differences stay in their asymptotic range
        du = 0.1 * (pack(smooth_state(shape, seed=seed)) - u)
        out.append((u, du, pp))
    return out


def test_rhs_tangent_matches_kernels_and_finite_differences(J):
    from jax_bssn_amd.tangent import rhs_tangent
    from jax_bssn_amd.time_evolve import rhs_packed
    for u, du, p in _cases(J):
        r, dr = rhs_tangent(u, du, p)
        assert _field_rel(r, rhs_packed(u, p)) <= 1e-10
        h = 1e-4
        fd = (rhs_packed(u + h * du, p) - rhs_packed(u - h * du, p)) / (2 * h)
        err = _field_rel(dr, fd)
        print("rhs tangent vs FD", f"{err:.1e}", flush=True)
        assert err <= 1e-6


def test_evolve_tangent_matches_finite_differences(J):
    from jax_bssn_amd.tangent import evolve_tangent
    from jax_bssn_amd.time_evolve import evolve_packed
    for u, du, p in _cases(J):
        s, ds = evolve_tangent(u, du, p, 3)
        ref = evolve_packed(u.clone(), p, 3)
        assert _field_rel(s, ref) <= 1e-12
        h = 1e-4
        fd = (evolve_packed(u + h * du, p, 3) - evolve_packed(u - h * du, p, 3)) / (2 * h)
        err = _field_rel(ds, fd)
        print("evolve tangent vs FD", f"{err:.1e}", flush=True)
        assert err <= 1e-6


def test_two_directions_equal_single_runs(J):
    from jax_bssn_amd.tangent import evolve_tangent, rhs_tangent
    for u, du, p in _cases(J):
        du2 = torch.roll(du, 3, dims=1) * 0.7 + 0.01 * du.flip(1)
        both = torch.stack([du, du2])
        r, dr = rhs_tangent(u, both, p)
        for l, d in enumerate((du, du2)):
            r1, dr1 = rhs_tangent(u, d, p)
            assert torch.equal(r, r1) or _field_rel(r, r1) <= 1e-13
            assert _field_rel(dr[l], dr1) <= 1e-13
        s, ds = evolve_tangent(u, both, p, 2)
        for l, d in enumerate((du, du2)):
            s1, ds1 = evolve_tangent(u, d, p, 2)
            assert _field_rel(s, s1) <= 1e-13
            err = _field_rel(ds[l], ds1)
            print("two directions vs one", l, f"{err:.1e}", flush=True)
            assert err <= 1e-13


def test_two_direction_tiles_are_deterministic_at_a_larger_grid(J):
    """A 32^3 Sommerfeld grid (interior, slab and edge face tiles all
    present): two runs of the two-direction evolution are bitwise equal, and
    equal the single-direction runs."""
    from jax_bssn_amd.tangent import evolve_tangent
    from jax_bssn_amd.variables import pack
    u = pack(smooth_state((32, 32, 32), seed=21))
    du = 0.1 * (pack(smooth_state((32, 32, 32), seed=23)) - u)
    du2 = torch.roll(du, 5, dims=2) * 0.5
    p = J.BSSNParameters(eta=2.0, g=0.75, nu=0.1, dx=0.05, dt=0.0125, zero_shift=0, gauge=1, xl_bc=1, xr_bc=1,
                         yl_bc=1, yr_bc=1, zl_bc=1, zr_bc=1, x_min=-0.8, y_min=-0.8, z_min=-0.8)
    both = torch.stack([du, du2])
    s_a, d_a = evolve_tangent(u, both, p, 2)
    s_b, d_b = evolve_tangent(u, both, p, 2)
    assert torch.equal(s_a, s_b) and torch.equal(d_a, d_b)
    for l, d in enumerate((du, du2)):
        _, d1 = evolve_tangent(u, d, p, 2)
        err = _field_rel(d_a[l], d1)
        print("32^3 two directions vs one", l, f"{err:.1e}", flush=True)
        assert err <= 1e-13


def _puncture_setup(J, n=24, L=4.0):
    from jax_bssn_amd.tangent import puncture_initial
    dx = 2 * L / n
    x = -L + (np.arange(n) + 0.5) * dx
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    c = np.random.default_rng(0).uniform(-1, 1, 3)
    initial = puncture_initial(X, Y, Z, width=1.0, lapse_centre=tuple(c))
    p = J.BSSNParameters(eta=2.0, kappa=0.0, nu=0.1, g=0.75, dx=dx, dt=0.25 * dx, zero_shift=0, gauge=1,
                         xl_bc=1, xr_bc=1, yl_bc=1, yr_bc=1, zl_bc=1, zr_bc=1, x_min=x[0], y_min=x[0], z_min=x[0])
    return initial, p


def _l2_forward(J, initial, p, theta, steps):
    from jax_bssn_amd.constraints import NORM_NAMES, constraints_packed, norms_packed
    from jax_bssn_amd.time_evolve import evolve_packed
    u, _ = initial(theta)
    evolve_packed(u, p, steps)
    cf = constraints_packed(u, p)
    return float(torch.sqrt(torch.mean(cf[0] * cf[0])))


def test_hamiltonian_gradient_matches_finite_differences_and_autograd(J):
    from jax_bssn_amd.tangent import HamiltonianL2AfterSteps
    initial, p = _puncture_setup(J)
    steps = 4
    theta0 = np.array([1.0, 1e-3])
    theta = torch.tensor(theta0, requires_grad=True)
    L = HamiltonianL2AfterSteps.apply(theta, initial, p, steps)
    L.backward()
    grad = theta.grad.numpy()
    assert abs(float(L) - _l2_forward(J, initial, p, theta0, steps)) <= 1e-12 * float(L)
    for k, h in ((0, 1e-4), (1, 1e-5)):
        tp, tm = theta0.copy(), theta0.copy()
        tp[k] += h
        tm[k] -= h
        fd = (_l2_forward(J, initial, p, tp, steps) - _l2_forward(J, initial, p, tm, steps)) / (2 * h)
        print("dL/dtheta", k, grad[k], "FD", fd, flush=True)
        assert abs(grad[k] - fd) <= 1e-6 * abs(fd)
