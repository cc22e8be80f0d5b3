"""CPU-side checks of the C ABI: the in-tree library loads, exports exactly the
functions include/bssn_amd.h declares, the ctypes mirrors have the C layout,
and invalid configurations fail before any GPU work (ValueError, as the
reference's validators raise)."""

import ctypes
import os
import re
import subprocess

import pytest

from jax_bssn_amd import _lib
from jax_bssn_amd.build import INCLUDE, LIB_PATH

HEADER = os.path.join(INCLUDE, "bssn_amd.h")


def _declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(bssn_[a-z0-9_]+)\s*\(", text)))


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(LIB_PATH):
        pytest.skip("library not built (run python -m jax_bssn_amd.build)")
    return _lib.load()


def test_header_and_binding_agree():
    assert _declared() == sorted(_lib.EXPORTED_SYMBOLS)


def test_library_exports_every_declared_symbol(lib):
    out = subprocess.run(["nm", "-D", "--defined-only", LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r"\bT (bssn_[a-z0-9_]+)$", out, flags=re.M))
    for name in _declared():
        assert name in exported, name
        assert hasattr(lib, name)


def test_struct_layout_matches_header():
    # doubles first (10 x 8), then 2 + 6 + 1 + 3 + 6 + 3 int32
    assert ctypes.sizeof(_lib.CParams) == 10 * 8 + 21 * 4 + 4  # padded to 8
    assert _lib.CParams.bc.offset == 88
    assert ctypes.sizeof(_lib.CGrid) == 12
    # 6 int32, 4 int64, then the 4 per-list face tile counts
    assert ctypes.sizeof(_lib.CBatchPlan) == 6 * 4 + 4 * 8 + 4 * 4
    assert _lib.CBatchPlan.total_cells.offset == 24 and _lib.CBatchPlan.face_list_tiles.offset == 56


def test_version_string(lib):
    assert b"gfx950" in lib.bssn_version()


def _call_rhs(lib, params, shape=(16, 16, 16)):
    return lib.bssn_rhs(ctypes.byref(_lib.c_params(params)), ctypes.byref(_lib.c_grid(shape)), 0,
                        None, None, None, None)


def test_invalid_configurations_raise_value_error(lib):
    from jax_bssn_amd.variables import BSSNParameters
    bad = [
        (BSSNParameters(dx=-1.0), "dx"),
        (BSSNParameters(dx=0.1, xl_bc=2), "boundary codes"),
        (BSSNParameters(dx=0.1, gauge=3), "gauge"),
        (BSSNParameters(dx=0.1, reduced_axes=(True, False, False), xl_bc=1), "reduced"),
        (BSSNParameters(dx=0.1, guard_cells=4, owned_shape=(16, 16, 16)), "owned range"),
    ]
    for params, word in bad:
        with pytest.raises(ValueError, match=word):
            _lib.check(_call_rhs(lib, params))
    with pytest.raises(ValueError, match="at least 8"):
        _lib.check(_call_rhs(lib, BSSNParameters(dx=0.1, zl_bc=1), shape=(16, 16, 6)))
    # a valid configuration reaches the pointer checks, still before any launch
    with pytest.raises(ValueError, match="null"):
        _lib.check(_call_rhs(lib, BSSNParameters(dx=0.1)))


def test_stage_argument_validation(lib):
    from jax_bssn_amd.variables import BSSNParameters
    p = ctypes.byref(_lib.c_params(BSSNParameters(dx=0.1)))
    g = ctypes.byref(_lib.c_grid((8, 8, 8)))
    with pytest.raises(ValueError, match="stage"):
        _lib.check(lib.bssn_rk4_stage(p, g, 0, 5, None, None, None, None, 0, None, None))
    buf = ctypes.c_void_p(0x1000)
    with pytest.raises(ValueError, match="alias"):
        _lib.check(lib.bssn_rk4_stage(p, g, 0, 2, buf, ctypes.c_void_p(0x2000), ctypes.c_void_p(0x3000),
                                      buf, 0, None, None))
    with pytest.raises(ValueError, match="precision"):
        _lib.check(lib.bssn_rhs(p, g, 7, buf, ctypes.c_void_p(0x4000), None, None))


def test_work_sizes(lib):
    from jax_bssn_amd.variables import BSSNParameters
    g = ctypes.byref(_lib.c_grid((8, 9, 10)))
    p0 = ctypes.byref(_lib.c_params(BSSNParameters(dx=0.1)))
    pk = ctypes.byref(_lib.c_params(BSSNParameters(dx=0.1, kappa=0.02)))
    n = 8 * 9 * 10
    assert lib.bssn_work_size(p0, g, 0) == 0
    assert lib.bssn_work_size(p0, g, 1) == 3 * 21 * n
    assert lib.bssn_work_size(pk, g, 0) == 30 * n
    assert lib.bssn_work_size(pk, g, 1) == 3 * 21 * n + 30 * n
    with pytest.raises(ValueError, match="kappa"):
        _lib.check(lib.bssn_rhs(pk, g, 0, ctypes.c_void_p(0x1000), ctypes.c_void_p(0x2000), None, None))


def test_tangent_argument_validation(lib):
    from jax_bssn_amd.variables import BSSNParameters
    g = ctypes.byref(_lib.c_grid((8, 9, 10)))
    n = 8 * 9 * 10
    assert lib.bssn_tangent_work_size(g, 1) == 3 * 2 * 21 * n
    assert lib.bssn_tangent_work_size(g, 2) == 3 * 3 * 21 * n
    with pytest.raises(ValueError, match="ndir"):
        _lib.check(lib.bssn_tangent_work_size(g, 3))
    p = ctypes.byref(_lib.c_params(BSSNParameters(dx=0.1)))
    buf = ctypes.c_void_p(0x1000)
    with pytest.raises(ValueError, match="ndir"):
        _lib.check(lib.bssn_rhs_tangent(p, g, 0, buf, buf, buf, buf, None))
    with pytest.raises(ValueError, match="ndir"):
        _lib.check(lib.bssn_evolve_tangent(p, g, 1, 3, buf, buf, buf, None))
    with pytest.raises(ValueError, match="null"):
        _lib.check(lib.bssn_hamiltonian_tangent(p, g, 2, buf, None, buf, buf, None))
    pk = ctypes.byref(_lib.c_params(BSSNParameters(dx=0.1, kappa=0.02)))
    with pytest.raises(NotImplementedError, match="kappa"):
        _lib.check(lib.bssn_evolve_tangent(pk, g, 1, 1, buf, buf, buf, None))


def test_ghost_fill_rejects_single_sample_linear_faces(lib):
    """A face tile owning one sample has no owned inward neighbour for the
    linear exterior slope; the fill refuses it before any launch."""
    thin = _lib.CTileLayout((ctypes.c_int32 * 3)(1, 2, 2), (ctypes.c_int32 * 3)(1, 4, 4), 1,
                            (ctypes.c_int32 * 3)(0, 0, 0))
    bcs = (ctypes.c_int32 * 6)(1, 1, 0, 0, 0, 0)
    buf = ctypes.c_void_p(0x1000)
    with pytest.raises(ValueError, match="two owned samples"):
        _lib.check(lib.bssn_tiled_ghost_fill(ctypes.byref(thin), bcs, 0, 1, buf, None))
    # the constant and preserve rules read only the face sample: accepted up
    # to the launch (not made here: the buffer is not a device pointer)
    periodic = (ctypes.c_int32 * 6)(0, 0, 0, 0, 0, 0)
    ok = _lib.CTileLayout((ctypes.c_int32 * 3)(1, 2, 2), (ctypes.c_int32 * 3)(2, 4, 4), 1,
                          (ctypes.c_int32 * 3)(0, 0, 0))
    with pytest.raises(ValueError, match="exterior"):
        _lib.check(lib.bssn_tiled_ghost_fill(ctypes.byref(ok), periodic, 3, 1, buf, None))


def test_packed_entry_points_reject_unsuitable_tensors():
    """The Python mirror hands bare pointers to the C ABI, so every packed
    entry point checks device, dtype, contiguity and shape first."""
    import torch

    from jax_bssn_amd import constraints as Cn
    from jax_bssn_amd import time_evolve as T
    from jax_bssn_amd.variables import BSSNParameters
    p = BSSNParameters(dx=0.1)
    host = torch.zeros((21, 8, 8, 8), dtype=torch.float64)
    with pytest.raises(ValueError, match="HIP device"):
        T.rhs_packed(host, p)
    with pytest.raises(ValueError, match="HIP device"):
        T.evolve_packed(host, p, 1)
    with pytest.raises(ValueError, match="HIP device"):
        T.project_packed(host)
    with pytest.raises(ValueError, match="HIP device"):
        Cn.constraints_packed(host, p)
    with pytest.raises(ValueError, match="HIP device"):
        T.Evolver(host, p)
    with pytest.raises(ValueError, match="torch tensor"):
        T.require_packed(host.numpy())
    meta = torch.empty((21, 8, 8, 8), dtype=torch.float64, device="meta")
    with pytest.raises(ValueError, match="HIP device"):
        T.require_packed(meta)
