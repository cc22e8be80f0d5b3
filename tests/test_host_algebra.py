"""CPU check of the kernels' per-point code against the FP64 oracle.

The host build (tests/hostcheck) runs the same rhs_point.h algebra, stencil
rows (one-sided Sommerfeld closures, MAD, upwind replacement), face RHS,
projection and stage-update emitter that the shell kernel runs on the GPU,
over every grid point.  Same bars as the GPU parity tests."""

import ctypes

import numpy as np
import pytest

from jax_bssn_amd import _lib
from oracle import bssn as OB
from tests.hostcheck import build as hb
from tests.parity_utils import rel_errors
from tests.test_gpu_parity import CASES, RHS_TOL, STEP_TOL

SYM = ((0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2))


@pytest.fixture(scope="module")
def host():
    if hb.stale():
        hb.build()
    lib = ctypes.CDLL(hb.LIB)
    lib.host_rhs.restype = ctypes.c_int
    lib.host_stage.restype = ctypes.c_int
    return lib


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


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def pack_matter(rho, S, Sij):
    m = np.empty((10,) + rho.shape)
    m[0] = rho
    m[1:4] = S
    for c, (i, j) in enumerate(SYM):
        m[4 + c] = Sij[i, j]
    return np.ascontiguousarray(m)


def host_rhs(host, v, p, matter=None):
    s = pack(v)
    out = np.zeros_like(s)
    err = ctypes.create_string_buffer(512)
    mp = _p(matter) if matter is not None else None
    rc = host.host_rhs(ctypes.byref(_lib.c_params(p)), ctypes.byref(_lib.c_grid(s.shape[1:])), _p(s), mp, _p(out),
                       err, 512)
    assert rc == 0, err.value
    return unpack(out)


def host_step(host, v, p):
    v = OB.enforce_algebraic_constraints(v)
    u0 = pack(v)
    ustar, ua, ub = np.zeros_like(u0), np.zeros_like(u0), np.zeros_like(u0)
    err = ctypes.create_string_buffer(512)
    cp, cg = ctypes.byref(_lib.c_params(p)), ctypes.byref(_lib.c_grid(u0.shape[1:]))
    for stage, us, out in ((1, u0, ua), (2, ua, ub), (3, ub, ua), (4, ua, ub)):
        rc = host.host_stage(cp, cg, stage, _p(us), _p(u0), _p(ustar), _p(out), 0, err, 512)
        assert rc == 0, err.value
    return unpack(ub)


@pytest.mark.parametrize("case", sorted(CASES))
def test_host_rhs_matches_oracle(host, case):
    v, p = CASES[case]()
    errs = rel_errors(host_rhs(host, v, p), OB.compute_bssn_rhs(v, p))
    print(case, {k: f"{e:.1e}" for k, e in errs.items()})
    assert max(errs.values()) <= RHS_TOL, errs


@pytest.mark.parametrize("case", ["gauge_wave", "sommerfeld_mixed", "mad_periodic", "kappa_periodic"])
def test_host_stage_update_matches_oracle_step(host, case):
    v, p = CASES[case]()
    errs = rel_errors(host_step(host, v, p), OB.rk4_step(v, p))
    print(case, {k: f"{e:.1e}" for k, e in errs.items()})
    assert max(errs.values()) <= STEP_TOL, errs


def matter_sources(shape, seed=21):
    rng = np.random.default_rng(seed)
    x = [np.arange(n) / n for n in shape]
    X, Y, Z = np.meshgrid(*x, indexing='ij')
    tp = 2 * np.pi

    def f(a):
        k = rng.integers(1, 3, 3)
        ph = rng.uniform(0, tp, 3)
        return a * (1.2 + np.sin(tp * k[0] * X + ph[0]) * np.cos(tp * k[1] * Y + ph[1]) * np.sin(tp * k[2] * Z + ph[2]))
    rho = f(0.02)
    S = np.stack([f(0.01) for _ in range(3)])
    Sij = np.zeros((3, 3) + tuple(shape))
    for i in range(3):
        for j in range(i, 3):
            Sij[i, j] = f(0.01)
            Sij[j, i] = Sij[i, j]
    return rho, S, Sij


@pytest.mark.parametrize("case", ["shift_1plog", "kappa_periodic", "sommerfeld_aligned"])
def test_host_rhs_with_matter_matches_oracle(host, case):
    v, p = CASES[case]()
    rho, S, Sij = matter_sources(v.conformal_factor.shape)
    ref = OB.compute_bssn_rhs_with_matter(v, p, rho, S, Sij)
    got = host_rhs(host, v, p, pack_matter(rho, S, Sij))
    errs = rel_errors(got, ref)
    print(case, {k: f"{e:.1e}" for k, e in errs.items()})
    assert max(errs.values()) <= RHS_TOL, errs


@pytest.mark.parametrize("case", ["gauge_wave", "shift_1plog", "sommerfeld_all"])
def test_host_rhs_f32_within_stated_error(host, case):
    from tests.test_gpu_parity import F32_RHS_TOL
    v, p = CASES[case]()
    s = pack(v)
    out = np.zeros_like(s)
    err = ctypes.create_string_buffer(512)
    rc = host.host_rhs_f32(ctypes.byref(_lib.c_params(p)), ctypes.byref(_lib.c_grid(s.shape[1:])), _p(s), _p(out),
                           err, 512)
    assert rc == 0, err.value
    errs = rel_errors(unpack(out), OB.compute_bssn_rhs(v, p))
    print(case, {k: f"{e:.1e}" for k, e in errs.items()})
    assert max(errs.values()) <= F32_RHS_TOL, errs


def test_xcd_remap_is_a_permutation(host):
    """Every launch of n workgroups (n up to 5000, and the 8 XCDs' share
    uneven) visits each of its n tiles exactly once after the per-XCD
    remapping."""
    assert host.host_xcd_remap_check(5000) == 0


@pytest.mark.parametrize("tile", [(4, 8, 8), (8, 8, 8), (2, 8, 16), (8, 8, 4)])
def test_tile_sets_partition_the_grid(host, tile):
    """The lean / closure tile launches cover every tile exactly once and
    split them by the closure predicate, for Sommerfeld on any subset of faces,
    tiled owned ranges and grids that are not a multiple of the tile."""
    from jax_bssn_amd.variables import BSSNParameters
    err = ctypes.create_string_buffer(512)
    rng = np.random.default_rng(5)
    configs = [((64, 64, 64), (1,) * 6, {}), ((64, 64, 64), (1, 0, 0, 1, 1, 1), {}),
               ((40, 24, 36), (0, 0, 1, 1, 0, 0), {}), ((16, 16, 16), (1,) * 6, {}),
               ((8, 8, 8), (1,) * 6, {}),
               ((38, 38, 38), (1,) * 6, dict(guard_cells=3, owned_shape=(32, 32, 32),
                                              owned_faces=(True, False, True, True, False, True)))]
    for _ in range(6):
        shape = tuple(int(v) for v in rng.integers(8, 70, 3))
        configs.append((shape, tuple(int(v) for v in rng.integers(0, 2, 6)), {}))
    for shape, bc, extra in configs:
        names = ("xl_bc", "xr_bc", "yl_bc", "yr_bc", "zl_bc", "zr_bc")
        p = BSSNParameters(dx=0.1, **dict(zip(names, bc)), **extra)
        rc = host.host_tile_check(ctypes.byref(_lib.c_params(p)), ctypes.byref(_lib.c_grid(shape)), *tile, err, 512)
        assert rc == 0, (shape, bc, extra, rc, err.value)
