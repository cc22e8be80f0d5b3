"""FP64 NumPy restatement of the reference's fixed-mesh-refinement step.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

Follows the frozen serial statement of the algorithm the reference keeps as
its own test oracle, tests/support/fmr_reference.py:
  * vertex-centred 2:1 children with four guard points (:23-67);
  * dense separable degree-5 Lagrange prolongation matrices (:84-112) and
    their x, y, z contraction (:123-127);
  * guard refresh that keeps the child's active data (:130-143);
  * injection of coincident child vertices (:146-160);
  * the stage-synchronous RK4 step: project, refill root -> leaf, RHS per
    level with the finest step and MAD weight q = (dx_finest / dx)^4 on
    coarser levels, final projection, injection leaf -> root with
    re-projection, refill (:256-310).
Tree hierarchies (``parent`` index per patch) follow reference
fmr/refinement.py:34-51 and fmr/tiled.py:39-48 (children carry periodic
face codes).
"""

import math
from typing import NamedTuple

import numpy as np

from . import bssn as OB


class FMRPatchSpec(NamedTuple):
    coarse_lo: tuple
    coarse_hi: tuple
    refinement_ratio: int = 2
    ghost_width: int = 4
    parent: object = None


class FMRHierarchySpec(NamedTuple):
    patches: tuple
    use_mad: bool = True


FACE_NAMES = ("xl_bc", "xr_bc", "yl_bc", "yr_bc", "zl_bc", "zr_bc")


def hierarchy_parents(hierarchy):
    """Reference fmr/refinement.py:34-51 (chain when ``parent`` is None)."""
    parents = [-1]
    for child, spec in enumerate(hierarchy.patches, start=1):
        parents.append(child - 1 if spec.parent is None else int(spec.parent))
    return tuple(parents)


def fine_active_shape(spec):
    """fmr_reference.py:47-50."""
    return tuple(spec.refinement_ratio * (h - l) + 1 for l, h in zip(spec.coarse_lo, spec.coarse_hi))


def fine_active_slice(spec):
    g = spec.ghost_width
    return tuple(slice(g, g + n) for n in fine_active_shape(spec))


def fine_coordinates(spec, dx_coarse, origin=(0.0, 0.0, 0.0)):
    """Coordinates of a padded child allocation (fmr_reference.py:62-81)."""
    g = spec.ghost_width
    axes = [origin[d] + (spec.coarse_lo[d] + (np.arange(fine_active_shape(spec)[d] + 2 * g) - g)
                         / spec.refinement_ratio) * dx_coarse for d in range(3)]
    return np.meshgrid(*axes, indexing="ij")


def interpolation_matrices(parent_shape, spec, parent_offset, periodic):
    """Dense prolongation matrices, one per axis (fmr_reference.py:84-112)."""
    mats = []
    for d in range(3):
        count = fine_active_shape(spec)[d] + 2 * spec.ghost_width
        M = np.zeros((count, parent_shape[d]))
        for i in range(count):
            t = parent_offset + spec.coarse_lo[d] + (i - spec.ghost_width) / spec.refinement_ratio
            nodes = [math.floor(t) + o for o in range(-2, 4)]
            if not periodic and (min(nodes) < 0 or max(nodes) >= parent_shape[d]):
                raise ValueError("child prolongation stencil exceeds parent allocation")
            for k, node in enumerate(nodes):
                w = math.prod((t - other) / (node - other) for m, other in enumerate(nodes) if m != k)
                M[i, node % parent_shape[d] if periodic else node] += w
        mats.append(M)
    return mats


def prolongate(parent, mats):
    """x, then y, then z contraction (fmr_reference.py:123-127)."""
    mx, my, mz = mats
    c = np.einsum("ia,...abc->...ibc", mx, parent)
    c = np.einsum("jb,...ibc->...ijc", my, c)
    return np.einsum("kc,...ijc->...ijk", mz, c)


def fill_fine_ghosts(parent_vars, child_vars, spec, mats):
    """Child guards from the parent, active data kept (fmr_reference.py:130-137)."""
    act = fine_active_slice(spec)
    out = []
    for p, c in zip(parent_vars, child_vars):
        s = prolongate(np.asarray(p), mats)
        s[(...,) + act] = np.asarray(c)[(...,) + act]
        out.append(s)
    return OB.BSSNVariables(*out)


def restrict_to_coarse(parent_vars, child_vars, spec, parent_ghost):
    """Injection of coincident active child vertices (fmr_reference.py:146-160)."""
    covered = tuple(slice(parent_ghost + l, parent_ghost + h + 1) for l, h in zip(spec.coarse_lo, spec.coarse_hi))
    r = spec.refinement_ratio
    out = []
    for p, c in zip(parent_vars, child_vars):
        p = np.array(p, copy=True)
        inj = np.asarray(c)[(...,) + fine_active_slice(spec)][..., ::r, ::r, ::r]
        p[(...,) + covered] = inj
        out.append(p)
    return OB.BSSNVariables(*out)


def evolution_parameters(parameters, hierarchy):
    """MAD weight from the finest spacing; periodic child faces (fmr/tiled.py:39-48)."""
    finest = min(p.dx for p in parameters)
    out = []
    for level, p in enumerate(parameters):
        ch = dict(mad_q=(finest / p.dx) ** 4 if hierarchy.use_mad else 1.0)
        if level:
            ch.update({n: 0 for n in FACE_NAMES})
        out.append(p._replace(**ch))
    return tuple(out)


def _matrices(states, hierarchy):
    parents = hierarchy_parents(hierarchy)
    mats = []
    for child, spec in enumerate(hierarchy.patches, start=1):
        par = parents[child]
        pshape = np.asarray(states[par].lapse).shape
        off = 0 if par == 0 else hierarchy.patches[par - 1].ghost_width
        mats.append(interpolation_matrices(pshape, spec, off, par == 0))
    return mats


def fill_hierarchy(states, hierarchy, mats=None):
    mats = _matrices(states, hierarchy) if mats is None else mats
    parents = hierarchy_parents(hierarchy)
    states = list(states)
    for child, spec in enumerate(hierarchy.patches, start=1):
        states[child] = fill_fine_ghosts(states[parents[child]], states[child], spec, mats[child - 1])
    return tuple(states)


def _add(v, k, s):
    return OB.BSSNVariables(*(np.asarray(a) + s * np.asarray(b) for a, b in zip(v, k)))


def fmr_rk4_step(states, parameters, hierarchy, return_stages=False):
    """One stage-synchronous RK4 step of the whole hierarchy
    (fmr_reference.py:256-310; ``parameters`` as given, normalised here)."""
    params = evolution_parameters(tuple(parameters), hierarchy)
    mats = _matrices(states, hierarchy)
    dt = params[0].dt

    def stage(sts):
        sts = tuple(OB.enforce_algebraic_constraints(s) for s in sts)
        sts = fill_hierarchy(sts, hierarchy, mats)
        return sts, tuple(OB.compute_bssn_rhs(s, p) for s, p in zip(sts, params))

    s0, k1 = stage(states)
    s1, k2 = stage(tuple(_add(v, k, dt / 2) for v, k in zip(s0, k1)))
    s2, k3 = stage(tuple(_add(v, k, dt / 2) for v, k in zip(s0, k2)))
    s3, k4 = stage(tuple(_add(v, k, dt) for v, k in zip(s0, k3)))
    result = [OB.BSSNVariables(*(np.asarray(x) + dt * (np.asarray(a) + 2 * np.asarray(b) + 2 * np.asarray(c)
                                                        + np.asarray(d)) / 6
                                 for x, a, b, c, d in zip(v, r1, r2, r3, r4)))
              for v, r1, r2, r3, r4 in zip(s0, k1, k2, k3, k4)]
    result = [OB.enforce_algebraic_constraints(s) for s in result]
    parents = hierarchy_parents(hierarchy)
    for child in range(len(result) - 1, 0, -1):
        par = parents[child]
        off = 0 if par == 0 else hierarchy.patches[par - 1].ghost_width
        result[par] = OB.enforce_algebraic_constraints(
            restrict_to_coarse(result[par], result[child], hierarchy.patches[child - 1], off))
    out = fill_hierarchy(tuple(result), hierarchy, mats)
    if return_stages:
        return out, (s0, s1, s2, s3)
    return out
