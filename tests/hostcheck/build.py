"""Compile tests/hostcheck/libhostcheck.so: the host build of the kernels'
per-point code (test infrastructure; the library never uses it)."""

import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
LIB = os.path.join(HERE, "libhostcheck.so")
SRCS = [os.path.join(HERE, "host_rhs.cpp"), os.path.join(HERE, "flop_count.cpp"), os.path.join(HERE, "adjoint_check.cpp")]
DEPS = SRCS + [os.path.join(ROOT, "jax_bssn_amd", "csrc", f) for f in
                ("bssn_common.h", "rhs_point.h", "stage_common.h", "stencil_providers.h", "tiles.h", "adjoint_point.h")] + \
       [os.path.join(ROOT, "include", "bssn_amd.h")]


def stale():
    if not os.path.exists(LIB):
        return True
    t = os.path.getmtime(LIB)
    return any(os.path.getmtime(d) > t for d in DEPS)


def build():
    rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
    cmd = [os.path.join(rocm, "bin", "hipcc"), "-O2", "-std=c++17", "-fPIC", "-shared", "--offload-arch=gfx950",
           f"-I{ROOT}/include", f"-I{ROOT}/jax_bssn_amd/csrc", *SRCS, "-o", LIB + ".tmp"]
    subprocess.run(cmd, check=True, capture_output=True)
    os.replace(LIB + ".tmp", LIB)
    return LIB
