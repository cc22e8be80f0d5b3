"""State / parameter containers and the packed device layout.

``BSSNVariables`` and ``BSSNParameters`` keep the reference names, field
order and defaults (reference bssn/variables.py:8-46) so code written against
the reference constructs them unchanged.  Fields may be NumPy arrays or torch
tensors in the reference layout: (3, 3, nx, ny, nz) tensors, (3, nx, ny, nz)
vectors, (nx, ny, nz) scalars, C order, z fastest.

On the GPU the state lives as one contiguous FP64 tensor of shape
(21, nx, ny, nz) -- the packed layout of include/bssn_amd.h: symmetric
tensors stored once (xx xy xz yy yz zz, read from the upper triangle like the
reference's Cholesky inverse, tensor_algebra.py:20-71) and unpacked
symmetrically.
"""

from typing import NamedTuple

import numpy as np
import torch

NCOMP = 21
SYM_PAIRS = ((0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2))


class BSSNVariables(NamedTuple):
    """Container for BSSN evolution variables (reference bssn/variables.py:8-16)."""
    conformal_metric: object      # gamma_tilde_ij (3, 3, ...)
    conformal_factor: object      # W
    traceless_K: object           # A_tilde_ij (3, 3, ...)
    trace_K: object               # K
    conformal_connection: object  # Gamma_tilde^i (3, ...)
    lapse: object                 # alpha
    shift: object                 # beta^i (3, ...)


class BSSNParameters(NamedTuple):
    """Parameters for BSSN evolution (reference bssn/variables.py:19-46)."""
    eta: float = 2.0
    kappa: float = 0.0
    nu: float = 0.002
    g: float = 0.75
    dx: float = 0.1
    dt: float = 0.001
    zero_shift: int = 0
    gauge: int = 0
    xl_bc: int = 0
    xr_bc: int = 0
    yl_bc: int = 0
    yr_bc: int = 0
    zl_bc: int = 0
    zr_bc: int = 0
    x_min: float = 0.0
    y_min: float = 0.0
    z_min: float = 0.0
    mad_q: float = 1.0
    guard_cells: int = 0
    owned_shape: tuple = None
    owned_faces: tuple = (True, True, True, True, True, True)
    reduced_axes: tuple = None


def get_boundary_codes(params, spatial_direction):
    """Left and right face codes of one axis (reference bssn/variables.py:49-59)."""
    if spatial_direction == 0:
        return params.xl_bc, params.xr_bc
    if spatial_direction == 1:
        return params.yl_bc, params.yr_bc
    return params.zl_bc, params.zr_bc


def device():
    if not torch.cuda.is_available():
        raise RuntimeError("jax_bssn_amd needs a HIP device (torch.cuda is unavailable); "
                           "there is no CPU path")
    return torch.device("cuda", torch.cuda.current_device())


def _as_tensor(x, dev):
    if isinstance(x, torch.Tensor):
        return x.to(device=dev, dtype=torch.float64)
    return torch.as_tensor(np.asarray(x, dtype=np.float64), device=dev)


def spatial_shape(v):
    return tuple(int(s) for s in np.shape(v.conformal_factor)[-3:])


def pack(v, out=None):
    """BSSNVariables (reference layout) -> (21, nx, ny, nz) FP64 device tensor."""
    dev = device()
    shape = spatial_shape(v)
    if out is None:
        out = torch.empty((NCOMP,) + shape, dtype=torch.float64, device=dev)
    g = _as_tensor(v.conformal_metric, dev)
    A = _as_tensor(v.traceless_K, dev)
    for c, (i, j) in enumerate(SYM_PAIRS):
        out[c].copy_(g[i, j])
        out[7 + c].copy_(A[i, j])
    out[6].copy_(_as_tensor(v.conformal_factor, dev))
    out[13].copy_(_as_tensor(v.trace_K, dev))
    out[14:17].copy_(_as_tensor(v.conformal_connection, dev))
    out[17].copy_(_as_tensor(v.lapse, dev))
    out[18:21].copy_(_as_tensor(v.shift, dev))
    return out


def _sym(packed, first):
    rows = []
    for i in range(3):
        row = []
        for j in range(3):
            row.append(packed[first + SYM_PAIRS.index((min(i, j), max(i, j)))])
        rows.append(torch.stack(row))
    return torch.stack(rows)


def unpack(packed, like=None):
    """(21, nx, ny, nz) device tensor -> BSSNVariables in the reference layout.

    ``like`` selects the output kind: a NumPy array gives NumPy fields, a
    torch tensor gives tensors on that tensor's device (default: device).
    """
    fields = BSSNVariables(
        conformal_metric=_sym(packed, 0),
        conformal_factor=packed[6].clone(),
        traceless_K=_sym(packed, 7),
        trace_K=packed[13].clone(),
        conformal_connection=packed[14:17].clone(),
        lapse=packed[17].clone(),
        shift=packed[18:21].clone(),
    )
    if like is None or isinstance(like, torch.Tensor):
        dev = like.device if isinstance(like, torch.Tensor) else packed.device
        return BSSNVariables(*(f.to(dev) for f in fields))
    return BSSNVariables(*(f.cpu().numpy() for f in fields))
