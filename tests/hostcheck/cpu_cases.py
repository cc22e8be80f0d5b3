"""Cases run on the CPU build of the library (libbssn_cpu.so, every kernel on
host threads under AddressSanitizer), each against the FP64 oracle.

Run as ``python -m tests.hostcheck.cpu_cases CASE ...`` in a process that
preloads the sanitizer runtime (cpu_build.asan_env()); an out-of-range
global or LDS index aborts the process with the sanitizer's report.  Prints
one line per case and exits non-zero on a mismatch.  Test infrastructure
only: the grids are small and tile-aligned where the tile kernels are
meant to run, so that every kernel family takes part:

* tiles_sommerfeld: interior tiles and the x, y, z-slab and edge/corner face
  tiles (24^3, all faces Sommerfeld), FP64 steps and the FP32-RHS RHS;
* shell: faces off the tile grid (the closure-shell point kernel);
* kappa_tiles / kappa_points: the constraint-damping term on the tile
  and on the point kernels;
* mad: the MAD blend tiles (periodic, mad_q < 1);
* cartoon_region / cartoon_points: the Cartoon centre plane on region tiles
  and on the point kernels (z-Sommerfeld, kappa = 0.02);
* tangent_tiles / tangent_points: one- and two-direction tangents against
  central differences of the CPU build's own RHS;
* batch: two uneven periodic instances and a Sommerfeld one in the batched
  launches (lists split into launches of 64 workgroups) against single-grid
  steps;
* poisoned_guards: NaN written into every sample beyond the stencils'
  reach of a tiled layout's owned points leaves the owned RHS unchanged.
"""

import ctypes
import os
import sys
import types

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

_lib = OB = OC = rel_errors = smooth_state = None


def _imports():
    """The package's ctypes structures without its torch-based __init__ (a
    bare package module stands in for it; only in the case process, never
    in a process that imports the package itself), the oracle and the
    parity helpers."""
    global _lib, OB, OC, rel_errors, smooth_state
    if "jax_bssn_amd" not in sys.modules:
        pkg = types.ModuleType("jax_bssn_amd")
        pkg.__path__ = [os.path.join(ROOT, "jax_bssn_amd")]
        sys.modules["jax_bssn_amd"] = pkg
    from jax_bssn_amd import _lib as lib_mod
    from oracle import bssn as ob
    from oracle import cartoon as oc
    from tests import parity_utils
    _lib, OB, OC = lib_mod, ob, oc
    rel_errors, smooth_state = parity_utils.rel_errors, parity_utils.smooth_state

LIB = os.path.join(HERE, "libbssn_cpu.so")
SYM = ((0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2))


def lib():
    L = ctypes.CDLL(LIB)
    for name, (res, args) in _lib._SIGNATURES.items():
        fn = getattr(L, name, None)
        if fn is not None:
            fn.restype, fn.argtypes = res, args
    return L


def pack(v):
    n = v.conformal_factor.shape
    s = np.empty((21,) + n)
    for c, (i, j) in enumerate(SYM):
        s[c] = v.conformal_metric[i, j]
        s[7 + c] = v.traceless_K[i, j]
    s[6] = v.conformal_factor
    s[13] = v.trace_K
    s[14:17] = v.conformal_connection
    s[17] = v.lapse
    s[18:21] = v.shift
    return np.ascontiguousarray(s)


def unpack(s):
    def sym(first):
        return np.stack([np.stack([s[first + SYM.index((min(i, j), max(i, j)))] for j in range(3)])
                         for i in range(3)])
    return OB.BSSNVariables(sym(0), s[6], sym(7), s[13], s[14:17], s[17], s[18:21])


def ptr(a):
    return ctypes.c_void_p(a.ctypes.data)


def check(L, rc):
    if rc != 0:
        raise RuntimeError(f"status {rc}: {L.bssn_last_error_string().decode()}")


def sommerfeld_params(dx=0.1, **kw):
    return OB.BSSNParameters(eta=2.0, g=0.75, nu=0.1, dx=dx, dt=0.25 * dx, zero_shift=0, gauge=1,
                             xl_bc=1, xr_bc=1, yl_bc=1, yr_bc=1, zl_bc=1, zr_bc=1,
                             x_min=-0.75, y_min=-0.8, z_min=-0.83)._replace(**kw)


def rhs(L, v, p, precision=0):
    s = pack(v)
    shape = s.shape[1:]
    cp, cg = _lib.c_params(p), _lib.c_grid(shape)
    out = np.zeros_like(s)
    work = np.zeros(max(1, L.bssn_work_size(ctypes.byref(cp), ctypes.byref(cg), 0)))
    check(L, L.bssn_rhs(ctypes.byref(cp), ctypes.byref(cg), precision, ptr(s), ptr(out), ptr(work), None))
    return out


def evolve(L, v, p, steps, precision=0):
    s = pack(v)
    shape = s.shape[1:]
    cp, cg = _lib.c_params(p), _lib.c_grid(shape)
    work = np.zeros(max(1, L.bssn_work_size(ctypes.byref(cp), ctypes.byref(cg), 1)))
    check(L, L.bssn_evolve(ctypes.byref(cp), ctypes.byref(cg), precision, steps, ptr(s), ptr(work), None))
    return s


def compare(name, got_packed, ref, tol):
    errs = rel_errors(unpack(got_packed), ref)
    worst = max(errs.values())
    print(f"{name}: max rel err {worst:.2e} (bar {tol:.0e})", flush=True)
    if not worst <= tol:
        raise AssertionError(f"{name}: {errs}")


def case_tiles_sommerfeld(L):
    v = smooth_state((24, 24, 24), seed=15)
    p = sommerfeld_params()
    compare("rhs f64", rhs(L, v, p), OB.compute_bssn_rhs(v, p), 1e-10)
    compare("rhs f32rhs", rhs(L, v, p, precision=1), OB.compute_bssn_rhs(v, p), 5e-4)
    ref = OB.rk4_step(OB.rk4_step(v, p), p)
    compare("two steps", evolve(L, v, p, 2), ref, 1e-12)


def case_shell(L):
    v = smooth_state((18, 16, 14), seed=9)
    p = sommerfeld_params(yl_bc=0, yr_bc=0)
    compare("rhs", rhs(L, v, p), OB.compute_bssn_rhs(v, p), 1e-10)
    compare("step", evolve(L, v, p, 1), OB.rk4_step(v, p), 1e-12)


def case_kappa_tiles(L):
    v = smooth_state((16, 16, 32), seed=14)
    p = sommerfeld_params(kappa=0.2)
    compare("rhs", rhs(L, v, p), OB.compute_bssn_rhs(v, p), 1e-10)
    compare("step", evolve(L, v, p, 1), OB.rk4_step(v, p), 1e-12)


def case_kappa_tiles_periodic(L):
    """The kappa term on the tiles of a periodic grid (every tile lean, the
    ordered prefetching staging), without and with matter sources."""
    from tests.test_host_algebra import matter_sources, pack_matter
    v = smooth_state((16, 16, 16), seed=22)
    p = OB.BSSNParameters(eta=0.3, g=0.75, nu=0.05, dx=1.0 / 16, dt=0.02, kappa=0.3)
    compare("rhs", rhs(L, v, p), OB.compute_bssn_rhs(v, p), 1e-10)
    rho, S, Sij = matter_sources(v.conformal_factor.shape)
    s = pack(v)
    cp, cg = _lib.c_params(p), _lib.c_grid(s.shape[1:])
    m = np.ascontiguousarray(pack_matter(rho, S, Sij))
    out = np.zeros_like(s)
    work = np.zeros(max(1, L.bssn_work_size(ctypes.byref(cp), ctypes.byref(cg), 0)))
    check(L, L.bssn_rhs_with_matter(ctypes.byref(cp), ctypes.byref(cg), 0, ptr(s), ptr(m), ptr(out), ptr(work),
                                    None))
    compare("rhs with matter", out, OB.compute_bssn_rhs_with_matter(v, p, rho, S, Sij), 1e-10)


def case_kappa_points(L):
    v = smooth_state((14, 15, 13), seed=21)
    p = sommerfeld_params(kappa=0.02, xl_bc=0, xr_bc=0)
    compare("rhs", rhs(L, v, p), OB.compute_bssn_rhs(v, p), 1e-10)


def case_mad(L):
    v = smooth_state((16, 16, 16), seed=3)
    p = OB.BSSNParameters(eta=0.3, g=0.75, nu=0.05, dx=1.0 / 16, dt=0.2 / 16, zero_shift=0, gauge=1,
                          mad_q=1.0 / 16.0)
    compare("rhs", rhs(L, v, p), OB.compute_bssn_rhs(v, p), 1e-10)
    compare("step", evolve(L, v, p, 1), OB.rk4_step(v, p), 1e-12)


def _cartoon_case(full_nx, kappa, z_somm=True):
    """The smooth axisymmetric state of the GPU Cartoon tests with z
    Sommerfeld (or periodic) faces and an O(1e-2) trace-free extrinsic
    curvature."""
    from tests.test_oracle_cartoon import _set_positive, full_and_compact
    x, y, z, dx, compact, cp = full_and_compact(full_nx)
    zb = 1 if z_somm else 0
    cp = cp._replace(nu=0.05, eta=1.0, g=0.75, kappa=kappa, dt=0.2 * dx, zl_bc=zb, zr_bc=zb)
    nr = compact.lapse.shape[0] - 4
    rho = (np.arange(nr) + 0.5) * dx
    R, Z = np.meshgrid(rho, z, indexing="ij")
    env = np.exp(-(R ** 2 + Z ** 2))
    A = np.zeros((3, 3) + R.shape)
    A[0, 0], A[1, 1], A[2, 2] = 0.02 * env, -0.01 * env, -0.01 * env
    A[0, 2] = A[2, 0] = 0.006 * R * Z * env
    compact = OC.fill_axisymmetric_ghosts(compact._replace(traceless_K=_set_positive(compact.traceless_K, A)))
    return compact, cp


def compact_rhs(L, s, p):
    """The call sequence of the Cartoon evolver's RHS: reconstruct the
    nine-plane support, the RHS on its centre plane, project back."""
    nx, _, nz = s.shape[1:]
    cg = _lib.c_grid((nx, 1, nz))
    cs = _lib.c_grid((nx, 9, nz))
    cp = _lib.c_params(p)
    sup = np.zeros((21, nx, 9, nz))
    sup_rhs = np.zeros_like(sup)
    work = np.zeros(max(1, L.bssn_work_size(ctypes.byref(cp), ctypes.byref(cs), 0)))
    lo = (ctypes.c_int32 * 3)(4, 4, 0)
    hi = (ctypes.c_int32 * 3)(nx, 5, nz)
    check(L, L.bssn_cartoon_axi_reconstruct(ctypes.byref(cg), p.dx, p.z_min, ptr(s), ptr(sup), None))
    check(L, L.bssn_rhs_region(ctypes.byref(cp), ctypes.byref(cs), 0, lo, hi, ptr(sup), ptr(sup_rhs), ptr(work),
                               None))
    out = np.zeros_like(s)
    check(L, L.bssn_cartoon_axi_project(ctypes.byref(cg), ptr(sup_rhs), ptr(out), None))
    return out


def case_cartoon_region(L):
    # full_nx = 32: 16 radial points and 32 z points on 16 x 1 x 16 region tiles
    v, p = _cartoon_case(32, 0.02)
    compare("compact rhs (region tiles)", compact_rhs(L, pack(v), p), OC.compute_axisymmetric_rhs(v, p), 1e-10)
    # periodic z: the region tiles' LDS halos wrap
    v, p = _cartoon_case(32, 0.02, z_somm=False)
    compare("compact rhs (region tiles, periodic z)", compact_rhs(L, pack(v), p), OC.compute_axisymmetric_rhs(v, p),
            1e-10)


def case_cartoon_points(L):
    # full_nx = 24: 12 radial points, off the region tile grid (point kernels)
    v, p = _cartoon_case(24, 0.02)
    compare("compact rhs (point kernels)", compact_rhs(L, pack(v), p), OC.compute_axisymmetric_rhs(v, p), 1e-10)


def _tangent_fd(L, v, p, nd, seed):
    """dRHS along nd random directions from bssn_rhs_tangent against
    central differences of the CPU build's own RHS, and each direction of an
    nd-direction call against a one-direction call."""
    s = pack(v)
    shape = s.shape[1:]
    cp, cg = _lib.c_params(p), _lib.c_grid(shape)
    rng = np.random.default_rng(seed)
    ds = np.ascontiguousarray(1e-2 * rng.standard_normal((nd,) + s.shape))
    out = np.zeros_like(s)
    dout = np.zeros_like(ds)
    check(L, L.bssn_rhs_tangent(ctypes.byref(cp), ctypes.byref(cg), nd, ptr(s), ptr(ds), ptr(out), ptr(dout), None))
    base = rhs(L, v, p)
    if np.max(np.abs(out - base)) > 1e-13 * np.max(np.abs(base)):
        raise AssertionError("value part of the tangent RHS differs from bssn_rhs")
    h = 1e-5
    for l in range(nd):
        fd = (rhs(L, unpack(s + h * ds[l]), p) - rhs(L, unpack(s - h * ds[l]), p)) / (2 * h)
        err = np.max(np.abs(dout[l] - fd)) / np.max(np.abs(fd))
        print(f"tangent nd={nd} dir {l}: rel diff to central differences {err:.2e} (bar 1e-6)", flush=True)
        if not err <= 1e-6:
            raise AssertionError(f"tangent direction {l}: {err}")
        if nd > 1:
            one = np.zeros_like(ds[l])
            dl = np.ascontiguousarray(ds[l])
            check(L, L.bssn_rhs_tangent(ctypes.byref(cp), ctypes.byref(cg), 1, ptr(s), ptr(dl), ptr(out), ptr(one),
                                        None))
            e1 = np.max(np.abs(dout[l] - one)) / np.max(np.abs(one))
            print(f"tangent nd={nd} dir {l} vs one direction: {e1:.2e} (bar 1e-13)", flush=True)
            if not e1 <= 1e-13:
                raise AssertionError(f"two-direction tangent {l} vs one direction: {e1}")


def case_tangent_tiles(L):
    # tile-aligned Sommerfeld faces: interior and face tiles of the dual tile kernel
    v = smooth_state((24, 24, 24), seed=31)
    p = sommerfeld_params()
    _tangent_fd(L, v, p, 2, 1)


def case_tangent_points(L):
    # faces off the tile grid: the one-direction point kernel, once per direction
    v = smooth_state((14, 16, 18), seed=32)
    p = sommerfeld_params()
    _tangent_fd(L, v, p, 2, 2)


def case_batch(L):
    """Two uneven periodic gauge-wave-like instances and a Sommerfeld one
    (lean and face tiles) advanced by the batched launches, each tile list
    covered by launches of at most 64 workgroups, against the same steps one
    grid at a time."""
    from jax_bssn_amd import _lib as LB
    shapes = ((16, 16, 16), (24, 16, 32), (24, 16, 32))
    states, params = [], []
    for k, sh in enumerate(shapes):
        v = smooth_state(sh, seed=40 + k, amp=0.3)
        if k < 2:
            params.append(OB.BSSNParameters(eta=0.0, g=0.0, nu=0.1, dx=1.0 / sh[0], dt=0.25 / sh[0], zero_shift=1,
                                            gauge=0))
        else:
            params.append(sommerfeld_params())
        states.append(pack(v))
    ref = [evolve(L, unpack(s), p, 2) for s, p in zip(states, params)]
    n = len(shapes)
    grids = (LB.CGrid * n)(*[LB.c_grid(sh) for sh in shapes])
    cps = (LB.CParams * n)(*[LB.c_params(p) for p in params])
    work = np.zeros(L.bssn_batch_work_size(grids, n))
    ptrs = (ctypes.c_void_p * n)(*[s.ctypes.data for s in states])
    plan = LB.CBatchPlan()
    check(L, L.bssn_batch_prepare(cps, grids, ptrs, n, 0, 0, 64, ptr(work), ctypes.byref(plan), None))
    for _ in range(2):
        check(L, L.bssn_batch_step(ctypes.byref(plan), ptr(work), None))
    for k, (s, r) in enumerate(zip(states, ref)):
        err = np.max(np.abs(s - r)) / np.max(np.abs(r))
        print(f"batch instance {k} {shapes[k]}: rel diff to single-grid steps {err:.2e} (bar 1e-13)", flush=True)
        if not err <= 1e-13:
            raise AssertionError(f"batch instance {k}: {err}")


def case_poisoned_guards(L):
    """A tile-local grid (guard_cells = 4 around 16^3 owned points, Sommerfeld
    faces owned on the low x, y and z sides): NaN in every guard sample more
    than three points from the owned box -- beyond every stencil's reach --
    must leave the RHS of the owned points bitwise unchanged (FP64 tile
    kernels with face tiles, and the FP32-RHS kernels)."""
    g, m = 4, 16
    n = m + 2 * g
    v = smooth_state((n, n, n), seed=51)
    p = sommerfeld_params(guard_cells=g, owned_shape=(m, m, m), owned_faces=(1, 0, 1, 0, 1, 0),
                          x_min=-0.75 - g * 0.1, y_min=-0.8 - g * 0.1, z_min=-0.83 - g * 0.1)
    s = pack(v)
    poisoned = s.copy()
    far = np.ones((n, n, n), bool)
    far[g - 3:g + m + 3, g - 3:g + m + 3, g - 3:g + m + 3] = False
    poisoned[:, far] = np.nan
    own = (slice(None), slice(g, g + m), slice(g, g + m), slice(g, g + m))
    for precision in (0, 1):
        a = rhs(L, unpack(s), p, precision)
        b = rhs(L, unpack(poisoned), p, precision)
        if not (np.isfinite(b[own]).all() and np.array_equal(a[own], b[own])):
            raise AssertionError(f"poisoned guards change the owned RHS (precision {precision})")
        print(f"poisoned guards (precision {precision}): owned RHS unchanged and finite", flush=True)


def case_constraints(L):
    """bssn_constraints against the oracle's compute_all_constraints on the
    tile kernels (periodic: every tile; aligned Sommerfeld: lean and face
    tiles) and the point kernel (unaligned faces, MAD)."""
    configs = {
        "periodic (tiles)": (smooth_state((12, 14, 16), seed=81),
                             OB.BSSNParameters(eta=0.3, g=0.75, nu=0.05, dx=1.0 / 12, dt=0.02)),
        "Sommerfeld aligned (lean and face tiles)": (smooth_state((24, 24, 24), seed=82), sommerfeld_params()),
        "Sommerfeld unaligned (points)": (smooth_state((14, 16, 18), seed=83), sommerfeld_params()),
        "MAD (points)": (smooth_state((12, 12, 12), seed=84),
                         OB.BSSNParameters(eta=0.3, g=0.75, nu=0.05, dx=1.0 / 12, dt=0.02, mad_q=0.25)),
    }
    for name, (v, p) in configs.items():
        s = pack(v)
        shape = s.shape[1:]
        out = np.zeros((9,) + shape)
        check(L, L.bssn_constraints(ctypes.byref(_lib.c_params(p)), ctypes.byref(_lib.c_grid(shape)), ptr(s),
                                    ptr(out), None))
        ref = OB.compute_all_constraints(v, p)
        fields = [ref.hamiltonian] + list(ref.momentum) + [ref.det_gamma, ref.trace_A] + list(ref.gamma_condition)
        worst = 0.0
        for k, b in enumerate(fields):
            b = np.asarray(b)
            err = np.abs(out[k] - b).max()
            scale = np.abs(b).max()
            if not (err <= 1e-10 * scale or err < 1e-13):
                raise AssertionError(f"constraints, {name}: field {k} error {err:.2e} (scale {scale:.2e})")
            worst = max(worst, err / max(scale, 1e-300) if err >= 1e-13 else 0.0)
        print(f"constraints, {name}: worst relative error {worst:.1e} (bar 1e-10)", flush=True)


def _vjp(L, s, p, rbar=None, hbar=None):
    shape = s.shape[1:]
    cp, cg = _lib.c_params(p), _lib.c_grid(shape)
    out = np.zeros_like(s)
    work = np.zeros(L.bssn_vjp_work_size(ctypes.byref(cg)))
    check(L, L.bssn_rhs_vjp(ctypes.byref(cp), ctypes.byref(cg), ptr(s), None if rbar is None else ptr(rbar),
                            None if hbar is None else ptr(hbar), ptr(out), 0, ptr(work), None))
    return out


def _dot_test(name, lhs, rhs_, tol=1e-12):
    err = abs(lhs - rhs_) / max(abs(lhs), abs(rhs_))
    print(f"{name}: <w, J v> = {lhs:.15e}, <J^T w, v> = {rhs_:.15e}, rel diff {err:.1e} (bar {tol:.0e})", flush=True)
    if not err <= tol:
        raise AssertionError(f"{name}: {err}")


def _adjoint_configs():
    return {
        "periodic 1+log": (smooth_state((12, 14, 16), seed=61),
                           OB.BSSNParameters(eta=0.3, g=0.75, nu=0.05, dx=1.0 / 12, dt=0.02, zero_shift=0, gauge=1)),
        "periodic harmonic, zero shift": (smooth_state((12, 12, 12), seed=62),
                                          OB.BSSNParameters(eta=0.0, g=0.0, nu=0.1, dx=1.0 / 12, dt=0.02,
                                                            zero_shift=1, gauge=0)),
        "Sommerfeld all faces": (smooth_state((16, 17, 18), seed=63), sommerfeld_params()),
        "Sommerfeld and periodic": (smooth_state((18, 16, 14), seed=64), sommerfeld_params(yl_bc=0, yr_bc=0)),
        # a plain interior [10, n - 10) on every axis: LDS tiles (partial ones included) and the gather shell
        "Sommerfeld, plain interior": (smooth_state((28, 26, 30), seed=66), sommerfeld_params()),
        "MAD": (smooth_state((14, 14, 14), seed=65),
                OB.BSSNParameters(eta=0.3, g=0.75, nu=0.05, dx=1.0 / 14, dt=0.02, zero_shift=0, gauge=1,
                                  mad_q=0.25)),
    }


def case_adjoint_rhs(L):
    """<w, J_F v> from the tangent kernels against <J_F^T w, v> from the
    adjoint kernels, for the RHS and for the Hamiltonian constraint."""
    rng = np.random.default_rng(7)
    for name, (v, p) in _adjoint_configs().items():
        s = pack(v)
        shape = s.shape[1:]
        cp, cg = _lib.c_params(p), _lib.c_grid(shape)
        dv = np.ascontiguousarray(1e-2 * rng.standard_normal(s.shape))
        w = np.ascontiguousarray(rng.standard_normal(s.shape))
        out, dout = np.zeros_like(s), np.zeros_like(s)
        check(L, L.bssn_rhs_tangent(ctypes.byref(cp), ctypes.byref(cg), 1, ptr(s), ptr(dv), ptr(out), ptr(dout), None))
        _dot_test(f"RHS, {name}", float(np.sum(w * dout)), float(np.sum(_vjp(L, s, p, rbar=w) * dv)))
        hb = np.ascontiguousarray(rng.standard_normal(shape))
        H, dH = np.zeros(shape), np.zeros(shape)
        check(L, L.bssn_hamiltonian_tangent(ctypes.byref(cp), ctypes.byref(cg), 1, ptr(s), ptr(dv), ptr(H), ptr(dH),
                                            None))
        _dot_test(f"H, {name}", float(np.sum(hb * dH)), float(np.sum(_vjp(L, s, p, hbar=hb) * dv)))


def case_adjoint_projection(L):
    rng = np.random.default_rng(8)
    v = smooth_state((10, 11, 12), seed=2)
    s = pack(v._replace(conformal_metric=v.conformal_metric * 1.1,
                        traceless_K=v.traceless_K + 0.05 * v.conformal_metric))
    cg = _lib.c_grid(s.shape[1:])
    dv = rng.standard_normal(s.shape)
    w = np.ascontiguousarray(rng.standard_normal(s.shape))

    def P(x):
        x = np.ascontiguousarray(x)
        check(L, L.bssn_enforce_constraints(ctypes.byref(cg), ptr(x), ptr(x), None))
        return x
    h = 1e-6
    lhs = float(np.sum(w * (P(s + h * dv) - P(s - h * dv)) / (2 * h)))
    wb = w.copy()
    check(L, L.bssn_project_vjp(ctypes.byref(cg), ptr(s), ptr(wb), None))
    _dot_test("projection (central differences)", lhs, float(np.sum(wb * dv)), tol=1e-8)


def _step_vjp(L, s, p, w):
    """J_step(s)^T w for one RK4 step, composed as jax_bssn_amd/adjoint.py
    composes it: recomputed stages, transposed RHS and projections in
    reverse order."""
    cg = _lib.c_grid(s.shape[1:])
    dt = p.dt

    def F(u):
        return rhs(L, unpack(u), p)

    def P(u):
        u = np.ascontiguousarray(u.copy())
        check(L, L.bssn_enforce_constraints(ctypes.byref(cg), ptr(u), ptr(u), None))
        return u

    def FT(u, kb):
        return _vjp(L, np.ascontiguousarray(u), p, rbar=np.ascontiguousarray(kb))

    def PT(u, ub):
        ub = np.ascontiguousarray(ub.copy())
        check(L, L.bssn_project_vjp(ctypes.byref(cg), ptr(np.ascontiguousarray(u)), ptr(ub), None))
        return ub

    u0 = P(s)
    k1 = F(u0)
    pre1 = u0 + 0.5 * dt * k1
    m1 = P(pre1)
    k2 = F(m1)
    pre2 = u0 + 0.5 * dt * k2
    m2 = P(pre2)
    k3 = F(m2)
    pre3 = u0 + dt * k3
    m3 = P(pre3)
    k4 = F(m3)
    pre_out = u0 + dt / 6.0 * (k1 + 2 * k2 + 2 * k3 + k4)
    pb = PT(pre_out, w)
    u0b = pb.copy()
    k1b, k2b, k3b, k4b = dt / 6.0 * pb, dt / 3.0 * pb, dt / 3.0 * pb, dt / 6.0 * pb
    b = PT(pre3, FT(m3, k4b))
    u0b += b
    k3b += dt * b
    b = PT(pre2, FT(m2, k3b))
    u0b += b
    k2b += 0.5 * dt * b
    b = PT(pre1, FT(m1, k2b))
    u0b += b
    k1b += 0.5 * dt * b
    u0b += FT(u0, k1b)
    return PT(s, u0b)


def case_adjoint_step(L):
    """The reverse sweep of one RK4 step against the tangent of the fused
    evolution step: <w, J_step v> = <J_step^T w, v>, Sommerfeld and periodic
    grids."""
    rng = np.random.default_rng(9)
    for name, (v, p) in _adjoint_configs().items():
        if p.mad_q != 1.0:
            continue
        s = pack(v)
        cp, cg = _lib.c_params(p), _lib.c_grid(s.shape[1:])
        dv = np.ascontiguousarray(1e-2 * rng.standard_normal(s.shape))
        w = rng.standard_normal(s.shape)
        u, du = s.copy(), dv.copy()
        work = np.zeros(L.bssn_tangent_work_size(ctypes.byref(cg), 1))
        check(L, L.bssn_evolve_tangent(ctypes.byref(cp), ctypes.byref(cg), 1, 1, ptr(u), ptr(du), ptr(work), None))
        _dot_test(f"RK4 step, {name}", float(np.sum(w * du)), float(np.sum(_step_vjp(L, s, p, w) * dv)))


def _complex_step(fn, v, dv, h=1e-30):
    """Directional derivative of the oracle map fn at v along dv, exact to
    rounding: Im fn(v + i h dv) / h (the oracle's arithmetic is analytic
    away from the upwind side select, whose derivative is zero)."""
    out = fn(OB.BSSNVariables(*[np.asarray(a, dtype=complex) + 1j * h * d for a, d in zip(v, dv)]))
    return np.imag(out) / h if isinstance(out, np.ndarray) else OB.BSSNVariables(*[np.imag(x) / h for x in out])


def case_adjoint_oracle(L):
    """The adjoint kernels against the derivative of the FP64 oracle itself
    (complex step through oracle/bssn.py, the stand-in for jax.grad of the
    reference): <w, J_oracle v> = <J^T w, v> for the RHS, the Hamiltonian
    constraint and one RK4 step.  The two sides differentiate the same
    discrete maps evaluated in different orders (observed <= 2e-15): 1e-12."""
    rng = np.random.default_rng(10)
    configs = _adjoint_configs()
    for name in ("periodic 1+log", "Sommerfeld all faces", "MAD", "Sommerfeld, plain interior"):
        v, p = configs[name]
        s = pack(v)
        dv = OB.BSSNVariables(*[1e-2 * rng.standard_normal(np.shape(x)) for x in v])
        sym = lambda t: 0.5 * (t + np.swapaxes(t, 0, 1))  # the packed state holds symmetric tensors
        dv = dv._replace(conformal_metric=sym(dv.conformal_metric), traceless_K=sym(dv.traceless_K))
        dvs = pack(dv)
        w = np.ascontiguousarray(rng.standard_normal(s.shape))
        jv = pack(_complex_step(lambda u: OB.compute_bssn_rhs(u, p), v, dv))
        _dot_test(f"RHS vs oracle, {name}", float(np.sum(w * jv)), float(np.sum(_vjp(L, s, p, rbar=w) * dvs)))
        hb = np.ascontiguousarray(rng.standard_normal(s.shape[1:]))
        jh = _complex_step(lambda u: OB.compute_hamiltonian_constraint(u, p), v, dv)
        _dot_test(f"H vs oracle, {name}", float(np.sum(hb * jh)), float(np.sum(_vjp(L, s, p, hbar=hb) * dvs)))
        if p.mad_q == 1.0:
            js = pack(_complex_step(lambda u: OB.rk4_step(u, p), v, dv))
            _dot_test(f"RK4 step vs oracle, {name}", float(np.sum(w * js)),
                      float(np.sum(_step_vjp(L, s, p, w) * dvs)))


CASES = {name[5:]: fn for name, fn in globals().items() if name.startswith("case_")}


def main(names):
    _imports()
    L = lib()
    bad = 0
    for name in names or sorted(CASES):
        print(f"== {name}", flush=True)
        try:
            CASES[name](L)
        except Exception as e:  # report and go on with the next case
            print(f"FAILED {name}: {e}", flush=True)
            bad += 1
    return 1 if bad else 0


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
