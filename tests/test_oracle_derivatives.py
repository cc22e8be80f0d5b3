"""Pins of the oracle's stencils by what mathematics fixes.

Each stencil is pinned by exact reproduction of polynomials of the degree its
width determines (which fixes the coefficients uniquely), by the Fourier
symbol of the periodic operators, and by formal convergence orders -- the
checks of reference tests/test_boundary_aware_derivatives.py,
tests/test_derivatives.py and tests/FMR/test_mad_derivative.py.
"""

import numpy as np
import pytest

from oracle.derivatives import (
    PERIODIC_BC,
    SOMMERFELD_BC,
    diff1_field,
    diff1_upwind_field,
    diff2_field,
    diff6_field,
)


def _line(n=24, dx=0.13, x0=-1.1):
    return x0 + dx * np.arange(n), dx


# ---- interior stencils: polynomial exactness (points away from the wrap) ----

def test_centred_d4_first_and_second_exact_on_quartic():
    x, dx = _line()
    f = 0.3 * x ** 4 - x ** 3 + 2 * x ** 2 - x + 0.5
    d1 = 1.2 * x ** 3 - 3 * x ** 2 + 4 * x - 1
    d2 = 3.6 * x ** 2 - 6 * x + 4
    inner = slice(4, -4)
    np.testing.assert_allclose(diff1_field(f, 0, dx)[inner], d1[inner], atol=1e-11)
    # D4 second derivative is exact for degree <= 5
    f5 = f + 0.1 * x ** 5
    d2_5 = d2 + 2.0 * x ** 3
    np.testing.assert_allclose(diff2_field(f5, 0, dx)[inner], d2_5[inner], atol=1e-10)


def test_centred_d6_exact_on_degree_six_with_mad_q_zero():
    x, dx = _line()
    f = 0.05 * x ** 6 - 0.2 * x ** 5 + x ** 3
    d1 = 0.3 * x ** 5 - x ** 4 + 3 * x ** 2
    d2 = 1.5 * x ** 4 - 4 * x ** 3 + 6 * x
    inner = slice(4, -4)
    np.testing.assert_allclose(diff1_field(f, 0, dx, mad_q=0.0)[inner], d1[inner], atol=1e-10)
    f7 = f + 0.01 * x ** 7
    d2_7 = d2 + 0.42 * x ** 5
    np.testing.assert_allclose(diff2_field(f7, 0, dx, mad_q=0.0)[inner], d2_7[inner], atol=1e-9)


def test_mad_blend_is_linear_in_q_and_q_one_is_d4():
    rng = np.random.default_rng(3)
    f = rng.normal(size=(17, 3, 5))
    d4 = diff1_field(f, 0, 0.2, mad_q=1.0)
    d6 = diff1_field(f, 0, 0.2, mad_q=0.0)
    q = 1.0 / 16.0
    np.testing.assert_allclose(diff1_field(f, 0, 0.2, mad_q=q), q * d4 + (1 - q) * d6,
                               atol=1e-13)
    old = (2 / 3 * (np.roll(f, -1, 0) - np.roll(f, 1, 0))
           - 1 / 12 * (np.roll(f, -2, 0) - np.roll(f, 2, 0))) / 0.2
    np.testing.assert_allclose(d4, old, rtol=0, atol=8 * np.finfo(float).eps * np.abs(old).max())


def test_upwind_stencils_exact_and_biased_the_reference_way():
    """c > 0 reads i-1..i+3 (forward), c < 0 reads i-3..i+1 (backward)."""
    x, dx = _line(n=30)
    f = 0.2 * x ** 4 - x ** 3 + x
    d1 = 0.8 * x ** 3 - 3 * x ** 2 + 1
    inner = slice(5, -5)
    for c in (2.0, -0.5, 0.0):
        np.testing.assert_allclose(diff1_upwind_field(f, c, 0, dx)[inner], d1[inner], atol=1e-11)
    # support: a spike at j only influences forward derivatives at j-3..j+1
    spike = np.zeros(30)
    spike[15] = 1.0
    fwd = diff1_upwind_field(spike, 1.0, 0, dx)
    bwd = diff1_upwind_field(spike, -1.0, 0, dx)
    assert set(np.nonzero(fwd)[0]) == {12, 13, 14, 15, 16}
    assert set(np.nonzero(bwd)[0]) == {14, 15, 16, 17, 18}
    # sixth-order upwind (MAD q = 0) exact on degree 6, support i-2..i+4
    f6 = f + 0.01 * x ** 6
    d6 = d1 + 0.06 * x ** 5
    np.testing.assert_allclose(diff1_upwind_field(f6, 1.0, 0, dx, mad_q=0.0)[inner],
                               d6[inner], atol=1e-10)
    assert set(np.nonzero(diff1_upwind_field(spike, 1.0, 0, dx, mad_q=0.0))[0]) == \
        set(range(11, 18))
    # the forward D4 stencil is dissipative for a rightward-travelling wave
    n = 64
    xx = 2 * np.pi * np.arange(n) / n
    wave = np.sin(3 * xx)
    dd = diff1_upwind_field(wave, 1.0, 0, 2 * np.pi / n) - diff1_field(wave, 0, 2 * np.pi / n)
    assert np.dot(dd, wave) < 0


def test_upwind_zero_coefficient_is_centred_and_selection_is_pointwise():
    rng = np.random.default_rng(814)
    f = rng.normal(size=19)
    coeff = np.resize(np.array([2.0, -0.5, 0.0]), 19)
    out = diff1_upwind_field(f, coeff, 0, 0.17)
    np.testing.assert_array_equal(out[coeff > 0], diff1_upwind_field(f, 1.0, 0, 0.17)[coeff > 0])
    np.testing.assert_array_equal(out[coeff < 0], diff1_upwind_field(f, -1.0, 0, 0.17)[coeff < 0])
    np.testing.assert_allclose(out[coeff == 0], diff1_field(f, 0, 0.17)[coeff == 0],
                               rtol=1e-14, atol=1e-14)


def test_periodic_fourier_symbols():
    """D4, D2 and D6 symbols on one Fourier mode (exact to roundoff)."""
    n, k = 40, 3
    h = 2 * np.pi / n
    x = h * np.arange(n)
    f = np.exp(1j * k * x)
    th = k * h
    sym1 = 1j * (4 / 3 * np.sin(th) - 1 / 6 * np.sin(2 * th)) / h
    sym2 = (-2 * np.cos(2 * th) + 32 * np.cos(th) - 30) / (12 * h * h)
    sym6 = (2 * np.cos(th) - 2) ** 3 / h ** 6
    for op, sym in ((lambda g: diff1_field(g, 0, h), sym1),
                    (lambda g: diff2_field(g, 0, h), sym2),
                    (lambda g: diff6_field(g, 0, h), sym6)):
        got = op(f.real) + 1j * op(f.imag)
        np.testing.assert_allclose(got, sym * f, atol=1e-9 * max(1, abs(sym)))


# ---- Sommerfeld closures: exactness and locality ----

def test_first_derivative_closures_exact_on_quartic_any_axis():
    x, dx = _line(n=12, dx=0.2, x0=-0.7)
    f1 = x ** 4
    ex = 4 * x ** 3
    idx = [0, 1, 10, 11]
    d = diff1_field(f1, 0, dx, SOMMERFELD_BC, SOMMERFELD_BC)
    np.testing.assert_allclose(d[idx], ex[idx], atol=1e-12)
    for axis in range(5):
        shape = [2] * 5
        shape[axis] = 12
        f = np.broadcast_to(f1.reshape([1] * axis + [12] + [1] * (4 - axis)), shape)
        d = diff1_field(f, axis, dx, SOMMERFELD_BC, SOMMERFELD_BC)
        np.testing.assert_allclose(np.take(d, idx, axis=axis),
                                   np.broadcast_to(ex[idx].reshape([1] * axis + [4] + [1] * (4 - axis)),
                                                   np.take(d, idx, axis=axis).shape), atol=1e-12)


def test_second_derivative_closures_exact_on_quintic():
    x, dx = _line(n=12, dx=0.2, x0=-0.7)
    f = x ** 5 - 0.4 * x ** 4 + 0.2 * x ** 2
    ex = 20 * x ** 3 - 4.8 * x ** 2 + 0.4
    d = diff2_field(f, 0, dx, SOMMERFELD_BC, SOMMERFELD_BC)
    idx = [0, 1, 10, 11]
    np.testing.assert_allclose(d[idx], ex[idx], atol=2e-11)


def test_sixth_difference_closures_exact_on_degree_seven():
    x, dx = _line(n=16, dx=0.2, x0=-1.3)
    f = x ** 7 - 2 * x ** 6 + x ** 3
    ex = 5040 * x - 1440
    d = diff6_field(f, 0, dx, SOMMERFELD_BC, SOMMERFELD_BC)
    idx = [0, 1, 2, 13, 14, 15]
    np.testing.assert_allclose(d[idx], ex[idx], rtol=1e-9, atol=2e-9 * np.abs(ex).max())
    # interior points of the array stay the periodic 7-point difference
    np.testing.assert_allclose(d[3:-3], ex[3:-3], rtol=1e-9, atol=2e-9 * np.abs(ex).max())


def test_closures_respect_interior_start_and_end():
    x, dx = _line(n=16, dx=0.2, x0=-1.3)
    f = x ** 4
    d = diff1_field(f, 0, dx, SOMMERFELD_BC, SOMMERFELD_BC, interior_start=3, interior_end=13)
    np.testing.assert_allclose(d[[3, 4, 11, 12]], 4 * x[[3, 4, 11, 12]] ** 3, atol=1e-11)


def test_upwind_sommerfeld_faces_take_centred_d4_and_never_wrap():
    x, dx = _line(n=12, dx=0.2, x0=-0.7)
    f = np.exp(x) + 0.1 * x ** 4
    coeff = np.resize(np.array([1.0, -1.0, 0.0]), 12)
    centred = diff1_field(f, 0, dx, SOMMERFELD_BC, SOMMERFELD_BC)
    up = diff1_upwind_field(f, coeff, 0, dx, SOMMERFELD_BC, SOMMERFELD_BC)
    np.testing.assert_array_equal(up[:3], centred[:3])
    np.testing.assert_array_equal(up[-3:], centred[-3:])
    np.testing.assert_array_equal(
        diff1_upwind_field(f, coeff, 0, dx, SOMMERFELD_BC, SOMMERFELD_BC, mad_q=0.0), up)
    g = np.sin(x) + 0.2 * x ** 2
    changed = g.copy()
    changed[-4:] = [1e6, -2e6, 3e6, -4e6]
    a = diff1_upwind_field(g, -1.0, 0, dx, SOMMERFELD_BC, PERIODIC_BC)
    b = diff1_upwind_field(changed, -1.0, 0, dx, SOMMERFELD_BC, PERIODIC_BC)
    np.testing.assert_array_equal(a[:3], b[:3])


# ---- convergence orders (tests/FMR/test_mad_derivative.py) ----

def _err(n, q, second=False):
    dx = 2 * np.pi / n
    x = np.arange(n) * dx
    f = np.sin(x)[:, None, None]
    if second:
        return np.sqrt(np.mean((diff2_field(f, 0, dx, mad_q=q) + f) ** 2))
    return np.sqrt(np.mean((diff1_field(f, 0, dx, mad_q=q) - np.cos(x)[:, None, None]) ** 2))


@pytest.mark.parametrize("second", [False, True])
def test_formal_orders_and_mad_leading_error(second):
    ns = (32, 64, 128, 256)
    o = lambda e: [np.log2(a / b) for a, b in zip(e[:-1], e[1:])]
    d4 = [_err(n, 1.0, second) for n in ns]
    mad = [_err(n, 1 / 16, second) for n in ns]
    d6 = [_err(n, 0.0, second) for n in ns]
    assert min(o(d4)[-2:]) > 3.9
    assert min(o(mad)[-2:]) > 3.8
    assert min(o(d6)[:2]) > 5.8
    ratios = [_err(c, 1 / 16, second) / _err(2 * c, 1.0, second) for c in (16, 256)]
    assert abs(ratios[-1] - 1) < 3e-3
    assert abs(ratios[-1] - 1) < abs(ratios[0] - 1)


def test_closure_orders():
    errs = []
    for n in (17, 33, 65):
        dx = 1.0 / (n - 1)
        x = dx * np.arange(n)
        f = np.exp(x)
        d = diff2_field(f, 0, dx, SOMMERFELD_BC, SOMMERFELD_BC)
        i = [0, 1, n - 2, n - 1]
        errs.append(np.max(np.abs(d[i] - f[i])))
    assert np.all(np.log2(np.array(errs[:-1]) / np.array(errs[1:])) > 3.7)
