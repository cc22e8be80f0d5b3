"""ctypes binding of ``libbssn_amd.so`` (the C ABI declared in include/bssn_amd.h).

The library is loaded from the package directory only; there is no fallback.
Status codes map to exceptions the way the reference surfaces errors:
invalid configuration -> ``ValueError``, HIP failure -> ``RuntimeError``.
"""

import ctypes
import os

from .build import LIB_PATH

NCOMP = 21
BSSN_OK = 0
BSSN_ERR_INVALID = -1
BSSN_ERR_HIP = -2
BSSN_ERR_UNSUPPORTED = -3
PREC_F64 = 0
PREC_F32RHS = 1
PART_TILES, PART_SHELL, PART_LEAN, PART_CLOSURE = 1, 2, 4, 8  # BSSN_PART_*


class CParams(ctypes.Structure):
    """Mirror of ``bssn_params``."""
    _fields_ = [
        ("eta", ctypes.c_double), ("kappa", ctypes.c_double), ("nu", ctypes.c_double),
        ("g", ctypes.c_double), ("dx", ctypes.c_double), ("dt", ctypes.c_double),
        ("x_min", ctypes.c_double), ("y_min", ctypes.c_double), ("z_min", ctypes.c_double),
        ("mad_q", ctypes.c_double),
        ("zero_shift", ctypes.c_int32), ("gauge", ctypes.c_int32),
        ("bc", ctypes.c_int32 * 6),
        ("guard_cells", ctypes.c_int32),
        ("owned_shape", ctypes.c_int32 * 3),
        ("owned_faces", ctypes.c_int32 * 6),
        ("reduced_axes", ctypes.c_int32 * 3),
    ]


class CGrid(ctypes.Structure):
    """Mirror of ``bssn_grid``."""
    _fields_ = [("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("nz", ctypes.c_int32)]


class CTileLayout(ctypes.Structure):
    """bssn_tile_layout of include/bssn_amd.h."""
    _fields_ = [("counts", ctypes.c_int32 * 3), ("tile_shape", ctypes.c_int32 * 3), ("guard_cells", ctypes.c_int32),
                ("upper_endpoints", ctypes.c_int32 * 3)]


class CBatchPlan(ctypes.Structure):
    """bssn_batch_plan of include/bssn_amd.h."""
    _fields_ = [("ninst", ctypes.c_int32), ("total_tiles", ctypes.c_int32), ("precision", ctypes.c_int32),
                ("tile_variant", ctypes.c_int32), ("grid_blocks", ctypes.c_int32), ("face_tiles", ctypes.c_int32),
                ("total_cells", ctypes.c_int64), ("desc_offset", ctypes.c_int64), ("prefix_offset", ctypes.c_int64),
                ("proj_offset", ctypes.c_int64), ("face_list_tiles", ctypes.c_int32 * 4)]


_SIGNATURES = {
    "bssn_last_error_string": (ctypes.c_char_p, []),
    "bssn_version": (ctypes.c_char_p, []),
    "bssn_work_size": (ctypes.c_int64, [ctypes.POINTER(CParams), ctypes.POINTER(CGrid), ctypes.c_int]),
    "bssn_rhs": (ctypes.c_int, [ctypes.POINTER(CParams), ctypes.POINTER(CGrid), ctypes.c_int,
                                ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "bssn_rhs_with_matter": (ctypes.c_int, [ctypes.POINTER(CParams), ctypes.POINTER(CGrid), ctypes.c_int,
                                            ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                            ctypes.c_void_p]),
    "bssn_enforce_constraints": (ctypes.c_int, [ctypes.POINTER(CGrid), ctypes.c_void_p,
                                                ctypes.c_void_p, ctypes.c_void_p]),
    "bssn_rk4_stage": (ctypes.c_int, [ctypes.POINTER(CParams), ctypes.POINTER(CGrid), ctypes.c_int,
                                      ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                      ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]),
    "bssn_evolve": (ctypes.c_int, [ctypes.POINTER(CParams), ctypes.POINTER(CGrid), ctypes.c_int,
                                   ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "bssn_constraints": (ctypes.c_int, [ctypes.POINTER(CParams), ctypes.POINTER(CGrid),
                                        ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "bssn_constraints_with_matter": (ctypes.c_int, [ctypes.POINTER(CParams), ctypes.POINTER(CGrid),
                                                    ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                                    ctypes.c_void_p]),
    "bssn_norm_partial_size": (ctypes.c_int64, [ctypes.POINTER(CGrid)]),
    "bssn_constraint_norms": (ctypes.c_int, [ctypes.POINTER(CGrid), ctypes.c_void_p, ctypes.c_void_p,
                                             ctypes.c_void_p, ctypes.c_void_p]),
    "bssn_constraint_norms_noalloc": (ctypes.c_int, [ctypes.POINTER(CGrid), ctypes.c_void_p, ctypes.c_void_p,
                                                     ctypes.c_void_p]),
    "bssn_fmr_prolong": (ctypes.c_int, [ctypes.POINTER(CGrid), ctypes.c_int32, ctypes.POINTER(ctypes.c_int32),
                                        ctypes.c_void_p, ctypes.POINTER(CGrid), ctypes.POINTER(ctypes.c_int32),
                                        ctypes.c_int32, ctypes.c_void_p, ctypes.c_void_p]),
    "bssn_fmr_restrict": (ctypes.c_int, [ctypes.POINTER(CGrid), ctypes.POINTER(ctypes.c_int32), ctypes.c_int32,
                                         ctypes.c_void_p, ctypes.POINTER(CGrid), ctypes.c_int32, ctypes.c_void_p,
                                         ctypes.c_void_p]),
    "bssn_rhs_region": (ctypes.c_int, [ctypes.POINTER(CParams), ctypes.POINTER(CGrid), ctypes.c_int,
                                       ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32),
                                       ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "bssn_rk4_combine": (ctypes.c_int, [ctypes.POINTER(CGrid), ctypes.c_double, ctypes.c_int, ctypes.c_void_p,
                                        ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                                        ctypes.c_void_p]),
    "bssn_cartoon_axi_fill_ghosts": (ctypes.c_int, [ctypes.POINTER(CGrid), ctypes.c_void_p, ctypes.c_void_p]),
    "bssn_cartoon_axi_reconstruct": (ctypes.c_int, [ctypes.POINTER(CGrid), ctypes.c_double, ctypes.c_double,
                                                    ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "bssn_cartoon_axi_project": (ctypes.c_int, [ctypes.POINTER(CGrid), ctypes.c_void_p, ctypes.c_void_p,
                                                ctypes.c_void_p]),
    "bssn_tiled_ghost_fill": (ctypes.c_int, [ctypes.POINTER(CTileLayout), ctypes.POINTER(ctypes.c_int32),
                                             ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]),
    "bssn_batch_work_size": (ctypes.c_int64, [ctypes.POINTER(CGrid), ctypes.c_int]),
    "bssn_batch_prepare": (ctypes.c_int, [ctypes.POINTER(CParams), ctypes.POINTER(CGrid),
                                          ctypes.POINTER(ctypes.c_void_p), ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                          ctypes.c_int, ctypes.c_void_p, ctypes.POINTER(CBatchPlan),
                                          ctypes.c_void_p]),
    "bssn_batch_step": (ctypes.c_int, [ctypes.POINTER(CBatchPlan), ctypes.c_void_p, ctypes.c_void_p]),
    "bssn_rhs_tangent": (ctypes.c_int, [ctypes.POINTER(CParams), ctypes.POINTER(CGrid), ctypes.c_int,
                                        ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                        ctypes.c_void_p]),
    "bssn_tangent_work_size": (ctypes.c_int64, [ctypes.POINTER(CGrid), ctypes.c_int]),
    "bssn_evolve_tangent": (ctypes.c_int, [ctypes.POINTER(CParams), ctypes.POINTER(CGrid), ctypes.c_int,
                                           ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                           ctypes.c_void_p]),
    "bssn_hamiltonian_tangent": (ctypes.c_int, [ctypes.POINTER(CParams), ctypes.POINTER(CGrid), ctypes.c_int,
                                                ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                                ctypes.c_void_p, ctypes.c_void_p]),
    "bssn_vjp_work_size": (ctypes.c_int64, [ctypes.POINTER(CGrid)]),
    "bssn_rhs_vjp": (ctypes.c_int, [ctypes.POINTER(CParams), ctypes.POINTER(CGrid), ctypes.c_void_p, ctypes.c_void_p,
                                    ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p,
                                    ctypes.c_void_p]),
    "bssn_project_vjp": (ctypes.c_int, [ctypes.POINTER(CGrid), ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "bssn_time_stage_part": (ctypes.c_int, [ctypes.POINTER(CParams), ctypes.POINTER(CGrid), ctypes.c_int,
                                            ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                            ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                                            ctypes.POINTER(ctypes.c_float), ctypes.c_void_p, ctypes.c_void_p]),
    "bssn_stage_points": (ctypes.c_int64, [ctypes.POINTER(CParams), ctypes.POINTER(CGrid), ctypes.c_int]),
}

EXPORTED_SYMBOLS = tuple(_SIGNATURES)

_lib = None


def load():
    """Load the in-tree library (raise if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it with `python -m jax_bssn_amd.build` "
            "(there is no CPU fallback)")
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in _SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(status):
    if status == BSSN_OK:
        return
    msg = load().bssn_last_error_string().decode(errors="replace")
    if status in (BSSN_ERR_INVALID,):
        raise ValueError(msg)
    if status == BSSN_ERR_UNSUPPORTED:
        raise NotImplementedError(msg)
    raise RuntimeError(msg)


def c_params(params):
    """BSSNParameters -> CParams (reference bssn/variables.py:19-46)."""
    p = CParams()
    for name in ("eta", "kappa", "nu", "g", "dx", "dt", "x_min", "y_min", "z_min", "mad_q"):
        setattr(p, name, float(getattr(params, name)))
    p.zero_shift = int(params.zero_shift)
    p.gauge = int(params.gauge)
    codes = (params.xl_bc, params.xr_bc, params.yl_bc, params.yr_bc, params.zl_bc, params.zr_bc)
    for i, c in enumerate(codes):
        p.bc[i] = int(c)
    p.guard_cells = int(params.guard_cells)
    if params.owned_shape is None:
        for i in range(3):
            p.owned_shape[i] = -1
    else:
        for i in range(3):
            p.owned_shape[i] = int(params.owned_shape[i])
    for i in range(6):
        p.owned_faces[i] = 1 if params.owned_faces[i] else 0
    if params.reduced_axes is None:
        for i in range(3):
            p.reduced_axes[i] = -1
    else:
        for i in range(3):
            p.reduced_axes[i] = 1 if params.reduced_axes[i] else 0
    return p


def work_size(params, shape, what):
    """Doubles of caller-owned scratch (bssn_work_size)."""
    n = load().bssn_work_size(ctypes.byref(c_params(params)), ctypes.byref(c_grid(shape)), int(what))
    if n < 0:
        check(int(n))
    return int(n)


def c_grid(shape):
    g = CGrid()
    g.nx, g.ny, g.nz = (int(s) for s in shape)
    return g
