"""Shared helpers of the parity tests: states, parameters, comparison metric."""

import numpy as np

from oracle import bssn as OB

GROUPS = (("conformal_metric", 0), ("conformal_factor", 1), ("traceless_K", 2), ("trace_K", 3),
          ("conformal_connection", 4), ("lapse", 5), ("shift", 6))


def rel_errors(a, b, floor=1e-300):
    """Per-field max-norm relative difference ||a - b||_inf / max(||b||_inf, floor).

    The norm of a tensor field is taken over all its components, so a
    component that is identically zero is judged against the scale of its
    tensor (and a field that is zero everywhere against ``floor``)."""
    out = {}
    for name, i in GROUPS:
        x = np.asarray(a[i], dtype=np.float64)
        y = np.asarray(b[i], dtype=np.float64)
        scale = max(float(np.max(np.abs(y))), floor)
        out[name] = float(np.max(np.abs(x - y))) / scale
    return out


def smooth_state(shape, seed=5, amp=1.0, lengths=(1.0, 1.0, 1.0)):
    """Smooth periodic non-solution state that exercises every RHS term."""
    rng = np.random.default_rng(seed)
    axes = [np.arange(n) / n for n in shape]
    X, Y, Z = np.meshgrid(*axes, indexing='ij')
    tp = 2 * np.pi

    def mode(a):
        k = rng.integers(1, 3, 3)
        ph = rng.uniform(0, tp, 3)
        return amp * a * (np.sin(tp * k[0] * X + ph[0]) * np.cos(tp * k[1] * Y + ph[1])
                          * np.sin(tp * k[2] * Z + ph[2]) + 0.3 * np.cos(tp * k[0] * X + ph[1]))
    g = np.zeros((3, 3) + tuple(shape))
    A = np.zeros_like(g)
    for i in range(3):
        for j in range(i, 3):
            g[i, j] = (1.0 if i == j else 0.0) + mode(0.05)
            g[j, i] = g[i, j]
            A[i, j] = mode(0.1)
            A[j, i] = A[i, j]
    v = OB.BSSNVariables(g, 1.0 + mode(0.1), A, mode(0.1), np.stack([mode(0.1) for _ in range(3)]),
                         1.0 + mode(0.1), np.stack([mode(0.1) for _ in range(3)]))
    return OB.enforce_algebraic_constraints(v)


def to_np(v):
    return OB.BSSNVariables(*(np.asarray(f.cpu().numpy() if hasattr(f, "cpu") else f, dtype=np.float64)
                              for f in v))
