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


def host_rhs(host, v, p):
    s = pack(v)
    out = np.zeros_like(s)
    err = ctypes.create_string_buffer(512)
    rc = host.host_rhs(ctypes.byref(_lib.c_params(p)), ctypes.byref(_lib.c_grid(s.shape[1:])), _p(s), _p(out),
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


@pytest.mark.parametrize("case", ["gauge_wave", "sommerfeld_mixed", "mad_periodic"])
def test_host_stage_update_matches_oracle_step(host, case):
    v, p = CASES[case]()
    errs = rel_errors(host_step(host, v, p), OB.rk4_step(v, p))
    print(case, {k: f"{e:.1e}" for k, e in errs.items()})
    assert max(errs.values()) <= STEP_TOL, errs
