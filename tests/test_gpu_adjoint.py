"""GPU tests of the reverse-mode adjoint (jax_bssn_amd/adjoint.py over
bssn_rhs_vjp / bssn_project_vjp), through the C ABI.

The adjoint of a linear map is pinned by the dot-product identity
<w, J v> = <J^T w, v> for random v, w, with J v from the forward-mode
tangent kernels (an independent implementation: dual numbers through the
evolution's own stage kernels) and from the FP64 oracle itself (complex
step through oracle/bssn.py, the stand-in for jax.grad of the reference).
Bars: 1e-12 relative for one RHS, H or step (both sides are exact
derivatives; only rounding separates them), the same over several
checkpointed steps, and 1e-8 for the gradient of the L2 Hamiltonian
constraint against the forward-mode gradient (the task's bar against
jax.grad)."""

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


def _dev(J, v):
    return J.pack(v).to("cuda").contiguous()


def _sommerfeld(shape, seed):
    v = smooth_state(shape, seed=seed)
    dx = 0.1
    p = OB.BSSNParameters(eta=2.0, g=0.75, nu=0.1, dx=dx, dt=0.25 * dx, zero_shift=0, gauge=1,
                          xl_bc=1, xr_bc=1, yl_bc=1, yr_bc=1, zl_bc=1, zr_bc=1, x_min=-0.75, y_min=-0.8,
                          z_min=-0.83)
    return v, p


def _periodic(shape, seed):
    v = smooth_state(shape, seed=seed)
    p = OB.BSSNParameters(eta=0.3, g=0.75, nu=0.05, dx=1.0 / shape[0], dt=0.2 / shape[0], zero_shift=0, gauge=1)
    return v, p


CASES = {
    # 24^3 tile-aligned: the tangent side runs the tile kernels with face tiles
    "sommerfeld_tiles": lambda: _sommerfeld((24, 24, 24), 71),
    "sommerfeld_points": lambda: _sommerfeld((14, 16, 18), 72),
    "periodic": lambda: _periodic((16, 16, 16), 73),
    # plain interior [10, n - 10): LDS tiles of the transpose, partial tiles and the gather shell
    "sommerfeld_interior": lambda: _sommerfeld((40, 36, 44), 74),
}


def _params(J, p):
    return J.BSSNParameters(**p._asdict())


def _check(name, lhs, rhs, tol):
    err = abs(lhs - rhs) / max(abs(lhs), abs(rhs))
    print(f"{name}: <w, J v> = {lhs:.15e}  <J^T w, v> = {rhs:.15e}  rel {err:.1e}", flush=True)
    assert err <= tol, (name, err)


@pytest.mark.parametrize("case", sorted(CASES))
def test_rhs_and_hamiltonian_adjoints_match_tangents(J, case):
    from jax_bssn_amd import adjoint as A
    from jax_bssn_amd import tangent as T
    v, p = CASES[case]()
    p = _params(J, p)
    print(f"[adjoint rhs] case={case}", flush=True)
    s = _dev(J, v)
    g = torch.Generator(device="cuda").manual_seed(1)
    dv = 1e-2 * torch.randn(s.shape, dtype=torch.float64, device="cuda", generator=g)
    w = torch.randn(s.shape, dtype=torch.float64, device="cuda", generator=g)
    ops = A._Ops(s, p)
    _, drhs = T.rhs_tangent(s, dv, p)
    _check(f"RHS {case}", float(torch.sum(w * drhs)), float(torch.sum(ops.FT(s, w) * dv)), 1e-12)
    hb = torch.randn(s.shape[1:], dtype=torch.float64, device="cuda", generator=g)
    H = torch.empty(s.shape[1:], dtype=torch.float64, device="cuda")
    dH = torch.empty_like(H)
    import ctypes
    from jax_bssn_amd import _lib
    from jax_bssn_amd.time_evolve import _ptr, stream_handle
    _lib.check(_lib.load().bssn_hamiltonian_tangent(ctypes.byref(_lib.c_params(p)),
                                                    ctypes.byref(_lib.c_grid(s.shape[1:])), 1, _ptr(s), _ptr(dv),
                                                    _ptr(H), _ptr(dH), stream_handle()))
    _check(f"H {case}", float(torch.sum(hb * dH)), float(torch.sum(ops.HT(s, hb.contiguous()) * dv)), 1e-12)


@pytest.mark.parametrize("case", ["sommerfeld_tiles", "periodic"])
def test_checkpointed_steps_adjoint_matches_tangent(J, case):
    """Six steps with a checkpoint every four (one full and one partial
    segment) against the tangent evolution of the same six steps."""
    from jax_bssn_amd import adjoint as A
    from jax_bssn_amd import tangent as T
    v, p = CASES[case]()
    p = _params(J, p)
    print(f"[adjoint steps] case={case}", flush=True)
    s = _dev(J, v)
    g = torch.Generator(device="cuda").manual_seed(2)
    dv = 1e-2 * torch.randn(s.shape, dtype=torch.float64, device="cuda", generator=g)
    w = torch.randn(s.shape, dtype=torch.float64, device="cuda", generator=g)
    _, du = T.evolve_tangent(s, dv, p, 6)
    ub, _ = A.evolve_vjp(s, 6, p, w, checkpoint_every=4)
    _check(f"6 steps {case}", float(torch.sum(w * du)), float(torch.sum(ub * dv)), 1e-12)
    # one step without checkpoints equals the single-step adjoint
    _, du1 = T.evolve_tangent(s, dv, p, 1)
    _check(f"1 step {case}", float(torch.sum(w * du1)), float(torch.sum(A.rk4_step_vjp(s, w, p) * dv)), 1e-12)


def _complex_step(fn, v, dv, h=1e-30):
    """Im fn(v + i h dv) / h: the oracle's exact directional derivative."""
    out = fn(OB.BSSNVariables(*[np.asarray(a, dtype=complex) + 1j * h * d for a, d in zip(v, dv)]))
    return OB.BSSNVariables(*[np.imag(x) / h for x in out])


@pytest.mark.parametrize("case", ["sommerfeld_tiles", "periodic", "sommerfeld_interior"])
def test_adjoint_matches_the_oracle_derivative(J, case):
    """<w, J_oracle v> with J_oracle v the complex-step derivative of the
    oracle's RHS and RK4 step against <J^T w, v> from the GPU adjoint."""
    from jax_bssn_amd import adjoint as A
    v, p = CASES[case]()
    print(f"[adjoint vs oracle] case={case}", flush=True)
    rng = np.random.default_rng(4)
    sym = lambda t: 0.5 * (t + np.swapaxes(t, 0, 1))
    dv = OB.BSSNVariables(*[1e-2 * rng.standard_normal(np.shape(x)) for x in v])
    dv = dv._replace(conformal_metric=sym(dv.conformal_metric), traceless_K=sym(dv.traceless_K))
    jv = _dev(J, _complex_step(lambda u: OB.compute_bssn_rhs(u, p), v, dv))
    js = _dev(J, _complex_step(lambda u: OB.rk4_step(u, p), v, dv))
    pj = _params(J, p)
    s, dvs = _dev(J, v), _dev(J, dv)
    g = torch.Generator(device="cuda").manual_seed(5)
    w = torch.randn(s.shape, dtype=torch.float64, device="cuda", generator=g)
    ops = A._Ops(s, pj)
    _check(f"RHS vs oracle {case}", float(torch.sum(w * jv)), float(torch.sum(ops.FT(s, w) * dvs)), 1e-12)
    _check(f"step vs oracle {case}", float(torch.sum(w * js)),
           float(torch.sum(A.rk4_step_vjp(s, w, pj, ops) * dvs)), 1e-12)


def test_reverse_gradient_matches_forward_mode_on_a_puncture(J):
    """Gradient of the L2 Hamiltonian constraint after 8 steps of a single
    puncture (Sommerfeld faces, 1+log, Gamma driver) with respect to the
    mass M and a Gaussian lapse perturbation eps: the reverse sweep
    (checkpoints every 4 steps) against the forward-mode tangents, <= 1e-8."""
    from jax_bssn_amd.adjoint import HamiltonianL2Reverse
    from jax_bssn_amd.tangent import HamiltonianL2AfterSteps, puncture_initial
    import bench
    n = 24
    L = 8.0
    dx = 2 * L / n
    x = -L + (np.arange(n) + 0.5) * dx
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    initial = puncture_initial(X, Y, Z, width=2.0, lapse_centre=(0.5, -0.3, 0.2))
    p = bench.puncture_params(J.BSSNParameters, dx, x[0])
    theta = torch.tensor([1.0, 1e-3], dtype=torch.float64, requires_grad=True)
    rev = HamiltonianL2Reverse.apply(theta, initial, p, 8, 4)
    rev.backward()
    g_rev = theta.grad.clone()
    theta2 = torch.tensor([1.0, 1e-3], dtype=torch.float64, requires_grad=True)
    fwd = HamiltonianL2AfterSteps.apply(theta2, initial, p, 8)
    fwd.backward()
    g_fwd = theta2.grad.clone()
    print("values", float(rev), float(fwd), "gradients", g_rev.tolist(), g_fwd.tolist(), flush=True)
    assert abs(float(rev) - float(fwd)) <= 1e-12 * abs(float(fwd))
    rel = (g_rev - g_fwd).abs() / g_fwd.abs()
    assert float(rel.max()) <= 1e-8, rel


def test_reverse_gradient_matches_the_oracle_complex_step(J):
    """The same objective and gradient as above, against the oracle itself:
    the puncture initial data, 8 RK4 steps, H and its L2 norm evaluated by
    the FP64 oracle on complex parameters (M + i h, eps + i h), whose
    imaginary parts are the exact derivatives -- the stand-in for jax.grad
    of the reference.  Value to 1e-10, gradient to 1e-8 (the task's bar)."""
    from jax_bssn_amd.adjoint import HamiltonianL2Reverse
    from jax_bssn_amd.tangent import puncture_initial
    import bench
    n, L, steps = 24, 8.0, 8
    dx = 2 * L / n
    x = -L + (np.arange(n) + 0.5) * dx
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    centre, width = (0.5, -0.3, 0.2), 2.0
    p = bench.puncture_params(J.BSSNParameters, dx, x[0])
    theta = torch.tensor([1.0, 1e-3], dtype=torch.float64, requires_grad=True)
    rev = HamiltonianL2Reverse.apply(theta, puncture_initial(X, Y, Z, width=width, lapse_centre=centre), p, steps, 4)
    rev.backward()
    g_rev = theta.grad.clone().numpy()
    op = OB.BSSNParameters(**p._asdict())

    def l2(M, eps):
        r = np.sqrt(X * X + Y * Y + Z * Z)
        W = (1.0 + M / (2.0 * r)) ** -2
        gauss = np.exp(-((X - centre[0]) ** 2 + (Y - centre[1]) ** 2 + (Z - centre[2]) ** 2) / width ** 2)
        zero = np.zeros((n, n, n), dtype=complex)
        g = np.zeros((3, 3, n, n, n), dtype=complex)
        g[0, 0] = g[1, 1] = g[2, 2] = 1.0
        v = OB.BSSNVariables(g, W + zero, np.zeros((3, 3, n, n, n), dtype=complex), zero.copy(),
                             np.zeros((3, n, n, n), dtype=complex), W + eps * gauss + zero,
                             np.zeros((3, n, n, n), dtype=complex))
        for _ in range(steps):
            v = OB.rk4_step(v, op)
        H = OB.compute_hamiltonian_constraint(v, op)
        return np.sqrt(np.mean(H * H))

    h = 1e-30
    print("[adjoint vs oracle complex step] puncture 24^3, 8 steps", flush=True)
    lM = l2(1.0 + 1j * h, 1e-3)
    le = l2(1.0, 1e-3 + 1j * h)
    g_ora = np.array([np.imag(lM) / h, np.imag(le) / h])
    print("value", float(rev), float(np.real(lM)), "gradient", g_rev.tolist(), g_ora.tolist(), flush=True)
    assert abs(float(rev) - float(np.real(lM))) <= 1e-10 * abs(float(np.real(lM)))
    rel = np.abs(g_rev - g_ora) / np.abs(g_ora)
    assert float(rel.max()) <= 1e-8, rel
