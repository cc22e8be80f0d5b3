"""FP64 NumPy restatement of the reference's tiled evolution on one device.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

Follows reference JAX_BSSN/parallel/:
  * layout geometry (layout.py:27-78): counts, tile_shape, guard_cells,
    upper_endpoints; allocated = tile_shape + 2 g;
  * the halo sweep (communication.py:36-106, _sweep_tile) with the linear /
    constant / preserve exterior rules;
  * tile-local parameters (evolution.py:19-38), owned masks (:41-50),
    project_owned (:95-102), tiled_rk4_step (:123-142) and
    tiled_constraint_norms (:145-191) reduced over owned samples.
Tiled arrays have the reference shape: components, then the tile grid,
then the allocated tile.
"""

from typing import NamedTuple

import numpy as np

from . import bssn as OB

FACE_NAMES = ("xl_bc", "xr_bc", "yl_bc", "yr_bc", "zl_bc", "zr_bc")


class TileLayout(NamedTuple):
    counts: tuple
    tile_shape: tuple
    guard_cells: int = 4
    upper_endpoints: tuple = (False, False, False)

    @property
    def active_shape(self):
        return tuple(w * c + int(e) for w, c, e in zip(self.tile_shape, self.counts, self.upper_endpoints))

    @property
    def allocated_shape(self):
        return tuple(w + 2 * self.guard_cells for w in self.tile_shape)


def _plane(axis, index):
    r = [slice(None)] * 3
    r[axis] = index
    return tuple(r)


def _sweep_tiles(tiles, bcs, layout, exterior):
    """communication.py:36-106 over a dict {tile index: (..., ax, ay, az)}."""
    g = layout.guard_cells
    counts = layout.counts
    for axis in range(3):
        width, count, endpoint = layout.tile_shape[axis], counts[axis], layout.upper_endpoints[axis]
        lower, upper = _plane(axis, slice(0, g)), _plane(axis, slice(-g, None))
        before = {k: v.copy() for k, v in tiles.items()}
        if count == 1:
            owned = width + int(endpoint)
            for k, t in tiles.items():
                n = t.shape[-3 + axis]
                nodes = g + (np.arange(n) - g) % owned
                tiles[k] = np.take(t, nodes, axis=t.ndim - 3 + axis)
        else:
            new = {}
            for k, t in before.items():
                lead = (Ellipsis,)
                lo_n = list(k)
                lo_n[axis] = (k[axis] - 1) % count
                hi_n = list(k)
                hi_n[axis] = (k[axis] + 1) % count
                src_lo = before[tuple(lo_n)]
                high_values = src_lo[lead + _plane(axis, slice(-2 * g, -g))]
                if endpoint and lo_n[axis] == count - 1:
                    high_values = src_lo[lead + _plane(axis, slice(width + 1, width + g + 1))]
                src_hi = before[tuple(hi_n)]
                upper_values = src_hi[lead + _plane(axis, slice(g, 2 * g))]
                if endpoint and k[axis] == count - 1:
                    upper_values = np.concatenate(
                        (t[lead + _plane(axis, slice(g + width, g + width + 1))],
                         upper_values[lead + _plane(axis, slice(0, g - 1))]), axis=t.ndim - 3 + axis)
                t = t.copy()
                t[lead + lower] = high_values
                t[lead + upper] = upper_values
                new[k] = t
            tiles = new
        for side in (0, 1):
            if not bcs[2 * axis + side]:
                continue
            for k, t in tiles.items():
                if k[axis] != (0 if side == 0 else count - 1):
                    continue
                e = int(endpoint and k[axis] == count - 1)
                target = lower if side == 0 else upper
                boundary = g if side == 0 else g + width - 1 + e
                nodes = np.arange(g) if side == 0 else np.arange(g + width, width + 2 * g)
                shape = [1, 1, 1]
                shape[axis] = g
                t = t.copy()
                if exterior == "preserve":
                    vals = before[k][(Ellipsis,) + target]
                else:
                    ax = t.ndim - 3 + axis
                    base = np.take(t, [boundary], axis=ax)
                    if exterior == "constant":
                        vals = np.broadcast_to(base, t[(Ellipsis,) + target].shape)
                    else:
                        inward = boundary + (1 if side == 0 else -1)
                        nb = np.take(t, [inward], axis=ax)
                        slope = nb - base if side == 0 else base - nb
                        vals = base + (nodes - boundary).reshape(shape) * slope
                t[(Ellipsis,) + target] = vals
                tiles[k] = t
    return tiles


def _split(field, layout):
    """(..., cx, cy, cz, ax, ay, az) -> {tile: (..., ax, ay, az)}."""
    c = layout.counts
    return {(i, j, k): np.array(field[..., i, j, k, :, :, :]) for i in range(c[0]) for j in range(c[1])
            for k in range(c[2])}


def _join(tiles, layout, like):
    out = np.empty_like(like)
    for (i, j, k), t in tiles.items():
        out[..., i, j, k, :, :, :] = t
    return out


def refresh_tiled_field(field, params, layout, exterior="linear"):
    bcs = tuple(int(getattr(params, n)) for n in FACE_NAMES)
    return _join(_sweep_tiles(_split(np.asarray(field), layout), bcs, layout, exterior), layout, np.asarray(field))


def tile_from_global(field, layout):
    """Owned samples of a guard-free global array into a zero-padded tiled
    array (initialization.py:41-83 without the refresh)."""
    field = np.asarray(field)
    lead = field.shape[:-3]
    g = layout.guard_cells
    out = np.zeros(lead + tuple(layout.counts) + layout.allocated_shape)
    for i in range(layout.counts[0]):
        for j in range(layout.counts[1]):
            for k in range(layout.counts[2]):
                idx = (i, j, k)
                owned = tuple(w + int(e and t == c - 1) for w, t, c, e in
                              zip(layout.tile_shape, idx, layout.counts, layout.upper_endpoints))
                src = tuple(slice(t * w, t * w + n) for t, w, n in zip(idx, layout.tile_shape, owned))
                dst = tuple(slice(g, g + n) for n in owned)
                out[(Ellipsis, i, j, k) + dst] = field[(Ellipsis,) + src]
    return out


def global_from_tiles(field, layout):
    """Owned samples of a tiled array back to the global array."""
    field = np.asarray(field)
    lead = field.shape[:-6]
    g = layout.guard_cells
    out = np.zeros(lead + layout.active_shape)
    for i in range(layout.counts[0]):
        for j in range(layout.counts[1]):
            for k in range(layout.counts[2]):
                idx = (i, j, k)
                owned = tuple(w + int(e and t == c - 1) for w, t, c, e in
                              zip(layout.tile_shape, idx, layout.counts, layout.upper_endpoints))
                dst = tuple(slice(t * w, t * w + n) for t, w, n in zip(idx, layout.tile_shape, owned))
                src = tuple(slice(g, g + n) for n in owned)
                out[(Ellipsis,) + dst] = field[(Ellipsis, i, j, k) + src]
    return out


def local_parameters(params, layout, idx):
    """evolution.py:19-38 for tile idx."""
    g = layout.guard_cells
    origins = (params.x_min, params.y_min, params.z_min)
    minima = tuple(o + (t * w - g) * params.dx for o, t, w in zip(origins, idx, layout.tile_shape))
    owned = tuple(w + int(e and t == c - 1) for w, t, c, e in
                  zip(layout.tile_shape, idx, layout.counts, layout.upper_endpoints))
    faces = tuple(f for t, c in zip(idx, layout.counts) for f in (t == 0, t == c - 1))
    reduced = params.reduced_axes if params.reduced_axes is not None else tuple(n == 1 for n in layout.active_shape)
    return params._replace(x_min=minima[0], y_min=minima[1], z_min=minima[2], guard_cells=g, owned_shape=owned,
                           owned_faces=faces, reduced_axes=reduced)


def owned_mask(shape, p):
    m = np.ones(shape, dtype=bool)
    for a, w in enumerate(p.owned_shape):
        n = np.arange(shape[a])
        dims = [1, 1, 1]
        dims[a] = shape[a]
        m &= ((n >= p.guard_cells) & (n < p.guard_cells + w)).reshape(dims)
    return m


def _tile_vars(state, idx):
    return OB.BSSNVariables(*(np.asarray(f)[(Ellipsis,) + idx + (slice(None),) * 3] for f in state))


def map_tiled(fn, state, params, layout):
    """evolution.py:53-71: fn(local_state, local_params) per tile, reassembled."""
    outs = {}
    for i in range(layout.counts[0]):
        for j in range(layout.counts[1]):
            for k in range(layout.counts[2]):
                outs[(i, j, k)] = fn(_tile_vars(state, (i, j, k)), local_parameters(params, layout, (i, j, k)))
    first = next(iter(outs.values()))
    fields = []
    for f_idx in range(len(first)):
        like = np.asarray(state[f_idx])
        out = np.empty_like(like)
        for (i, j, k), v in outs.items():
            out[..., i, j, k, :, :, :] = v[f_idx]
        fields.append(out)
    return type(first)(*fields)


def refresh_bssn_halos(state, params, layout):
    return OB.BSSNVariables(*(refresh_tiled_field(f, params, layout) for f in state))


def project_owned(state, params, layout):
    def proj(local, p):
        pr = OB.enforce_algebraic_constraints(local)
        m = owned_mask(np.asarray(local.lapse).shape, p)
        return OB.BSSNVariables(*(np.where(m, a, b) for a, b in zip(pr, local)))
    return map_tiled(proj, state, params, layout)


def prepare_tiled_state(state, params, layout):
    return refresh_bssn_halos(project_owned(state, params, layout), params, layout)


def _add(v, k, s):
    return OB.BSSNVariables(*(np.asarray(a) + s * np.asarray(b) for a, b in zip(v, k)))


def tiled_rk4_step(state, params, layout):
    """evolution.py:123-142."""
    dt = params.dt
    rhs = lambda s: map_tiled(OB.compute_bssn_rhs, s, params, layout)
    state = prepare_tiled_state(state, params, layout)
    k1 = rhs(state)
    k2 = rhs(prepare_tiled_state(_add(state, k1, dt / 2), params, layout))
    k3 = rhs(prepare_tiled_state(_add(state, k2, dt / 2), params, layout))
    k4 = rhs(prepare_tiled_state(_add(state, k3, dt), params, layout))
    res = OB.BSSNVariables(*(np.asarray(q) + dt / 6 * (np.asarray(a) + 2 * np.asarray(b) + 2 * np.asarray(c)
                                                      + np.asarray(d)) for q, a, b, c, d in zip(state, k1, k2, k3, k4)))
    return prepare_tiled_state(res, params, layout)
