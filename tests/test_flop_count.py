"""The algorithmic FP64 operation count bench.py reports against the FP64
vector peak: recounted here by the host build of the per-point code
(tests/hostcheck/flop_count.cpp) with a counting scalar type.

The count is that of the algorithm, not of the data: it is the same at
every interior point and for either sign of the shift (the forward and
backward upwind rows have five non-zero taps each)."""

import ctypes

import numpy as np
import pytest

import bench
from jax_bssn_amd import _lib
from jax_bssn_amd.variables import BSSNParameters
from tests.hostcheck import build as hb


@pytest.fixture(scope="module")
def host():
    if hb.stale():
        hb.build()
    lib = ctypes.CDLL(hb.LIB)
    lib.host_flop_count.restype = ctypes.c_int
    return lib


def _count(host, state, point, params):
    n = state.shape[1:]
    out = (ctypes.c_longlong * 10)()
    err = ctypes.create_string_buffer(256)
    rc = host.host_flop_count(ctypes.byref(_lib.c_params(params)), ctypes.byref(_lib.c_grid(n)),
                              state.ctypes.data_as(ctypes.c_void_p), *point, out, err, 256)
    assert rc == 0, err.value
    return list(out)


def _generic_state(n, seed, shift_sign=None):
    rng = np.random.default_rng(seed)
    s = 0.01 * rng.standard_normal((21, n, n, n))
    s[[0, 3, 5]] += 1.0
    s[6] += 0.7
    s[17] += 0.6
    if shift_sign is not None:
        s[18:21] = shift_sign * (0.05 + np.abs(s[18:21]))
    return np.ascontiguousarray(s)


def test_counts_match_the_bench_constants(host):
    n = 16
    params = bench.puncture_params(BSSNParameters, 1.0 / n, -0.5)
    c = _count(host, _generic_state(n, 1), (8, 8, 8), params)
    rhs, proj = sum(c[:5]), sum(c[5:])
    assert rhs == bench.FLOPS_RHS, c[:5]
    assert proj == bench.FLOPS_PROJECTION, c[5:]
    # one reciprocal per inverse metric, one cube root per projection
    assert c[2] == 2 and c[7] == 2 and c[9] == 1
    assert bench.F_ALG == 4 * rhs + 5 * proj + 21 * 14


@pytest.mark.parametrize("sign", [1.0, -1.0])
def test_count_is_independent_of_point_and_shift_sign(host, sign):
    n = 16
    params = bench.puncture_params(BSSNParameters, 1.0 / n, -0.5)
    state = _generic_state(n, 2, shift_sign=sign)
    counts = {tuple(_count(host, state, pt, params)) for pt in ((5, 6, 7), (8, 8, 8), (10, 4, 9))}
    assert counts == {tuple(_count(host, _generic_state(n, 1), (8, 8, 8), params))}
