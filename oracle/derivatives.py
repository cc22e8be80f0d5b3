"""FP64 NumPy restatement of the reference finite-difference operators.

TEST INFRASTRUCTURE ONLY.  This module is the CPU oracle against which the HIP
kernels are checked.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import it; the product path never
does.

Each function restates the array program of the reference file cited in its
docstring with ``np.roll`` in place of ``jnp.roll`` and ``if`` in place of
``lax.cond`` (parameters are concrete Python numbers here).

Pins (tests/test_oracle_derivatives.py): closed-form periodic D1/D6 stencils,
exact polynomial reproduction of every Sommerfeld closure (x^4 for D1, quintic
for D2, x^7 for D6), the upwind sign convention and its Sommerfeld-face
replacement, MAD q=1 == D4 and q=0 == D6, formal convergence orders.
"""

import numpy as np

PERIODIC_BC = 0
SOMMERFELD_BC = 1


def _end(field, direction, interior_end):
    return field.shape[direction] if interior_end is None else interior_end


def _shift(field, offset, direction):
    """``field[i + offset]`` with periodic wrap (``jnp.roll(field, -offset)``)."""
    return np.roll(field, -offset, axis=direction)


def diff1_field(field, direction, dx, left_bc=0, right_bc=0, mad_q=1.0,
                interior_start=0, interior_end=None):
    """Centred first derivative (reference evolution/derivatives.py:42-162).

    D4 = (2/3 (f1 - b1) - 1/12 (f2 - b2)) / dx; for ``mad_q != 1`` blended
    with the centred D6.  A Sommerfeld code on either face forces D4 on the
    axis; an active face replaces its outer two points by the one-sided
    closures of derivatives.py:99-153.
    """
    field = np.asarray(field, dtype=np.result_type(field, np.float64))
    end = _end(field, direction, interior_end)
    f1, f2, f3 = (_shift(field, k, direction) for k in (1, 2, 3))
    b1, b2, b3 = (_shift(field, -k, direction) for k in (1, 2, 3))
    d4 = 2 / 3 * (f1 - b1) / dx + -1 / 12 * (f2 - b2) / dx
    if mad_q == 1.0 or left_bc == SOMMERFELD_BC or right_bc == SOMMERFELD_BC:
        out = d4
    else:
        d6 = (3 / 4 * (f1 - b1) - 3 / 20 * (f2 - b2) + 1 / 60 * (f3 - b3)) / dx
        out = mad_q * d4 + (1.0 - mad_q) * d6
    out = np.moveaxis(np.array(out), direction, -1)
    src = np.moveaxis(field, direction, -1)
    s = interior_start
    if left_bc == SOMMERFELD_BC:
        left_0 = (-25.0 / 12.0 * src[..., s] + 4.0 * src[..., s + 1]
                  - 3.0 * src[..., s + 2] + 4.0 / 3.0 * src[..., s + 3]
                  - 1.0 / 4.0 * src[..., s + 4]) / dx
        left_1 = (-1.0 / 4.0 * src[..., s] - 5.0 / 6.0 * src[..., s + 1]
                  + 3.0 / 2.0 * src[..., s + 2] - 1.0 / 2.0 * src[..., s + 3]
                  + 1.0 / 12.0 * src[..., s + 4]) / dx
        out[..., s] = left_0
        out[..., s + 1] = left_1
    if right_bc == SOMMERFELD_BC:
        e = end
        right_1 = (-1.0 / 12.0 * src[..., e - 5] + 1.0 / 2.0 * src[..., e - 4]
                   - 3.0 / 2.0 * src[..., e - 3] + 5.0 / 6.0 * src[..., e - 2]
                   + 1.0 / 4.0 * src[..., e - 1]) / dx
        right_0 = (1.0 / 4.0 * src[..., e - 5] - 4.0 / 3.0 * src[..., e - 4]
                   + 3.0 * src[..., e - 3] - 4.0 * src[..., e - 2]
                   + 25.0 / 12.0 * src[..., e - 1]) / dx
        out[..., e - 2] = right_1
        out[..., e - 1] = right_0
    return np.moveaxis(out, -1, direction)


def diff1_upwind_field(field, rhs_coefficient, direction, dx, left_bc=0,
                       right_bc=0, mad_q=1.0, interior_start=0,
                       interior_end=None):
    """Upwind-biased first derivative (reference derivatives.py:166-338).

    ``rhs_coefficient > 0`` selects the forward-biased stencil, ``< 0`` the
    backward-biased one and ``== 0`` the centred D4.  On a Sommerfeld axis
    D4 is forced and the outer three points of each active face take the
    boundary-aware centred D4.
    """
    field = np.asarray(field, dtype=np.result_type(field, np.float64))
    end = _end(field, direction, interior_end)
    f1, f2, f3 = (_shift(field, k, direction) for k in (1, 2, 3))
    b1, b2, b3 = (_shift(field, -k, direction) for k in (1, 2, 3))
    d4_forward = (-3.0 * b1 - 10.0 * field + 18.0 * f1 - 6.0 * f2 + f3) / (12.0 * dx)
    d4_backward = (-b3 + 6.0 * b2 - 18.0 * b1 + 10.0 * field + 3.0 * f1) / (12.0 * dx)
    d4_centered = (2.0 / 3.0 * (f1 - b1) - 1.0 / 12.0 * (f2 - b2)) / dx
    physical_axis = left_bc == SOMMERFELD_BC or right_bc == SOMMERFELD_BC
    if mad_q == 1.0 or physical_axis:
        forward, backward, centered = d4_forward, d4_backward, d4_centered
    else:
        f4 = _shift(field, 4, direction)
        b4 = _shift(field, -4, direction)
        d6_forward = (2.0 * b2 - 24.0 * b1 - 35.0 * field + 80.0 * f1
                      - 30.0 * f2 + 8.0 * f3 - f4) / (60.0 * dx)
        d6_backward = (b4 - 8.0 * b3 + 30.0 * b2 - 80.0 * b1 + 35.0 * field
                       + 24.0 * f1 - 2.0 * f2) / (60.0 * dx)
        d6_centered = (3.0 / 4.0 * (f1 - b1) - 3.0 / 20.0 * (f2 - b2)
                       + 1.0 / 60.0 * (f3 - b3)) / dx
        forward = mad_q * d4_forward + (1.0 - mad_q) * d6_forward
        backward = mad_q * d4_backward + (1.0 - mad_q) * d6_backward
        centered = mad_q * d4_centered + (1.0 - mad_q) * d6_centered
    # the side is chosen by the real part (piecewise: its derivative is zero)
    coeff = np.real(np.asarray(rhs_coefficient)).astype(np.float64)
    derivative = np.where(coeff > 0.0, forward,
                          np.where(coeff < 0.0, backward, centered))
    derivative = np.array(np.broadcast_to(derivative, field.shape))
    if physical_axis:
        centred = diff1_field(field, direction, dx, left_bc, right_bc, mad_q=1.0,
                              interior_start=interior_start, interior_end=end)
        width = min(3, field.shape[direction])
        sel = np.moveaxis(derivative, direction, -1)
        cen = np.moveaxis(centred, direction, -1)
        if left_bc == SOMMERFELD_BC:
            sel[..., interior_start:interior_start + width] = \
                cen[..., interior_start:interior_start + width]
        if right_bc == SOMMERFELD_BC:
            sel[..., end - width:end] = cen[..., end - width:end]
        derivative = np.moveaxis(sel, -1, direction)
    return derivative


def diff2_field(field, direction, dx, left_bc=0, right_bc=0, mad_q=1.0,
                interior_start=0, interior_end=None):
    """Pure second derivative (reference derivatives.py:341-461)."""
    field = np.asarray(field, dtype=np.result_type(field, np.float64))
    end = _end(field, direction, interior_end)
    f1, f2, f3 = (_shift(field, k, direction) for k in (1, 2, 3))
    b1, b2, b3 = (_shift(field, -k, direction) for k in (1, 2, 3))
    d4 = (-f2 + 16.0 * f1 - 30.0 * field + 16.0 * b1 - b2) / (12.0 * dx ** 2)
    if mad_q == 1.0 or left_bc == SOMMERFELD_BC or right_bc == SOMMERFELD_BC:
        out = d4
    else:
        d6 = (1.0 / 90.0 * f3 - 3.0 / 20.0 * f2 + 3.0 / 2.0 * f1
              - 49.0 / 18.0 * field + 3.0 / 2.0 * b1 - 3.0 / 20.0 * b2
              + 1.0 / 90.0 * b3) / dx ** 2
        out = mad_q * d4 + (1.0 - mad_q) * d6
    out = np.moveaxis(np.array(out), direction, -1)
    src = np.moveaxis(field, direction, -1)
    s = interior_start
    if left_bc == SOMMERFELD_BC:
        left_0 = (15.0 / 4.0 * src[..., s] - 77.0 / 6.0 * src[..., s + 1]
                  + 107.0 / 6.0 * src[..., s + 2] - 13.0 * src[..., s + 3]
                  + 61.0 / 12.0 * src[..., s + 4] - 5.0 / 6.0 * src[..., s + 5]) / dx ** 2
        left_1 = (5.0 / 6.0 * src[..., s] - 5.0 / 4.0 * src[..., s + 1]
                  - 1.0 / 3.0 * src[..., s + 2] + 7.0 / 6.0 * src[..., s + 3]
                  - 1.0 / 2.0 * src[..., s + 4] + 1.0 / 12.0 * src[..., s + 5]) / dx ** 2
        out[..., s] = left_0
        out[..., s + 1] = left_1
    if right_bc == SOMMERFELD_BC:
        e = end
        right_1 = (1.0 / 12.0 * src[..., e - 6] - 1.0 / 2.0 * src[..., e - 5]
                   + 7.0 / 6.0 * src[..., e - 4] - 1.0 / 3.0 * src[..., e - 3]
                   - 5.0 / 4.0 * src[..., e - 2] + 5.0 / 6.0 * src[..., e - 1]) / dx ** 2
        right_0 = (-5.0 / 6.0 * src[..., e - 6] + 61.0 / 12.0 * src[..., e - 5]
                   - 13.0 * src[..., e - 4] + 107.0 / 6.0 * src[..., e - 3]
                   - 77.0 / 6.0 * src[..., e - 2] + 15.0 / 4.0 * src[..., e - 1]) / dx ** 2
        out[..., e - 2] = right_1
        out[..., e - 1] = right_0
    return np.moveaxis(out, -1, direction)


# One-sided sixth-difference closures (reference derivatives.py:511-590),
# rows for the points start, start+1, start+2 over taps start..start+7.
D6_LEFT_CLOSURES = (
    (4.0, -27.0, 78.0, -125.0, 120.0, -69.0, 22.0, -3.0),
    (3.0, -20.0, 57.0, -90.0, 85.0, -48.0, 15.0, -2.0),
    (2.0, -13.0, 36.0, -55.0, 50.0, -27.0, 8.0, -1.0),
)


def diff6_field(field, direction, dx, left_bc=0, right_bc=0, interior_start=0,
                interior_end=None):
    """Sixth difference for Kreiss-Oliger (reference derivatives.py:463-599)."""
    field = np.asarray(field, dtype=np.result_type(field, np.float64))
    end = _end(field, direction, interior_end)
    f1, f2, f3 = (_shift(field, k, direction) for k in (1, 2, 3))
    b1, b2, b3 = (_shift(field, -k, direction) for k in (1, 2, 3))
    d6 = (f3 + (-6) * f2 + 15 * f1 + (-20) * field + 15 * b1 + (-6) * b2 + b3) / (dx ** 6)
    out = np.moveaxis(np.array(d6), direction, -1)
    src = np.moveaxis(field, direction, -1)
    s = interior_start
    if left_bc == SOMMERFELD_BC:
        rows = []
        for row in D6_LEFT_CLOSURES:
            acc = row[0] * src[..., s]
            for t in range(1, 8):
                acc = acc + row[t] * src[..., s + t]
            rows.append(acc / dx ** 6)
        for r, value in enumerate(rows):
            out[..., s + r] = value
    if right_bc == SOMMERFELD_BC:
        e = end
        rows = []
        for r, row in enumerate(D6_LEFT_CLOSURES):
            # mirror: point e-1-r uses taps e-1 .. e-8 with the same weights
            acc = row[0] * src[..., e - 1]
            for t in range(1, 8):
                acc = acc + row[t] * src[..., e - 1 - t]
            rows.append(acc / dx ** 6)
        for r, value in enumerate(rows):
            out[..., e - 1 - r] = value
    return np.moveaxis(out, -1, direction)


# --------------------------------------------------------------------------
# Parameter-aware wrappers (reference evolution/spatial_derivatives.py:14-92)
# --------------------------------------------------------------------------

def boundary_codes(params, d):
    """(left, right) face codes of spatial axis ``d`` (variables.py:49-59)."""
    return ((params.xl_bc, params.xr_bc), (params.yl_bc, params.yr_bc),
            (params.zl_bc, params.zr_bc))[d]


def boundary_data(shape, d, params):
    """Closure data of spatial axis ``d`` (spatial_derivatives.py:18-41)."""
    left_bc, right_bc = boundary_codes(params, d)
    mad_q = 1.0 if (left_bc == SOMMERFELD_BC or right_bc == SOMMERFELD_BC) else params.mad_q
    left_bc = left_bc if params.owned_faces[2 * d] else 0
    right_bc = right_bc if params.owned_faces[2 * d + 1] else 0
    start = params.guard_cells
    count = shape[d] if params.owned_shape is None else params.owned_shape[d]
    if params.reduced_axes is not None:
        reduced = bool(params.reduced_axes[d])
    elif params.owned_shape is not None:
        reduced = count == 1
    else:
        reduced = False
    return left_bc, right_bc, start, start + count, mad_q, reduced


def _spatial(field, direction):
    return direction % field.ndim - (field.ndim - 3)


def diff1_physical(field, direction, params):
    d = _spatial(field, direction)
    lb, rb, s, e, q, reduced = boundary_data(field.shape[-3:], d, params)
    out = diff1_field(field, direction, params.dx, lb, rb, mad_q=q,
                      interior_start=s, interior_end=e)
    return np.zeros_like(out) if reduced else out


def diff2_physical(field, direction, params):
    d = _spatial(field, direction)
    lb, rb, s, e, q, reduced = boundary_data(field.shape[-3:], d, params)
    out = diff2_field(field, direction, params.dx, lb, rb, mad_q=q,
                      interior_start=s, interior_end=e)
    return np.zeros_like(out) if reduced else out


def diff1_upwind_physical(field, rhs_coefficient, direction, params):
    d = _spatial(field, direction)
    lb, rb, s, e, q, reduced = boundary_data(field.shape[-3:], d, params)
    out = diff1_upwind_field(field, rhs_coefficient, direction, params.dx, lb,
                             rb, mad_q=q, interior_start=s, interior_end=e)
    return np.zeros_like(out) if reduced else out


def ko_dissipation(field, params):
    """nu/64 h^5 sum_d D6_d field (spatial_derivatives.py:81-92)."""
    terms = []
    for d in range(3):
        h = params.dx
        lb, rb, s, e, _, reduced = boundary_data(field.shape[-3:], d, params)
        term = h ** 5 * diff6_field(field, field.ndim - 3 + d, h, lb, rb,
                                    interior_start=s, interior_end=e)
        terms.append(np.zeros_like(term) if reduced else term)
    return params.nu / 64 * (terms[0] + terms[1] + terms[2])
