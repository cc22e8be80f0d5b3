"""Build the in-tree HIP library ``jax_bssn_amd/libbssn_amd.so`` for gfx950.

Run ``python -m jax_bssn_amd.build`` (or ``__graft_entry__.build()``).  The
library is compiled with hipcc straight from ``jax_bssn_amd/csrc`` and the
public header in ``include/``; there is no other build system.
"""

import os
import subprocess
import sys

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG_DIR)
CSRC = os.path.join(PKG_DIR, "csrc")
INCLUDE = os.path.join(ROOT, "include")
LIB_NAME = "libbssn_amd.so"
LIB_PATH = os.path.join(PKG_DIR, LIB_NAME)
ARCH = os.environ.get("BSSN_OFFLOAD_ARCH", "gfx950")

SOURCES = ["bssn_amd.hip", "stage_f64.hip", "stage_f32.hip", "transfer.hip", "cartoon.hip", "halo.hip", "batch.hip", "tangent.hip", "tangent_tile1.hip", "tangent_tile2.hip", "adjoint.hip", "record_f64.hip"]

HIPCC_FLAGS = [
    "-O3",
    "-std=c++17",
    f"--offload-arch={ARCH}",
    "-fPIC",
    "-shared",
    # keep the compiler from forming packed FP32 instructions
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def hipcc():
    rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
    path = os.path.join(rocm, "bin", "hipcc")
    return path if os.path.exists(path) else "hipcc"


def build(verbose=True, extra_flags=()):
    """Compile every HIP source into LIB_PATH (always recompiles): one object
    per source, compiled in parallel, then one shared-library link."""
    import tempfile
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    tmp = LIB_PATH + ".tmp"
    compile_flags = [f for f in HIPCC_FLAGS if f != "-shared"]
    with tempfile.TemporaryDirectory(prefix="bssn_amd_build_") as objdir:
        objs, procs = [], []
        for src in srcs:
            obj = os.path.join(objdir, os.path.basename(src) + ".o")
            cmd = [hipcc(), *compile_flags, *extra_flags, f"-I{INCLUDE}", f"-I{CSRC}", "-c", src, "-o", obj]
            if verbose:
                print(" ".join(cmd), flush=True)
            procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)))
            objs.append(obj)
        failed = []
        for src, proc in procs:
            out, _ = proc.communicate()
            if proc.returncode != 0:
                failed.append((src, out))
        if failed:
            for src, out in failed:
                sys.stderr.write(f"--- {src}\n{out}")
            raise RuntimeError(f"hipcc failed on {[os.path.basename(s) for s, _ in failed]}")
        link = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp]
        if verbose:
            print(" ".join(link), flush=True)
        res = subprocess.run(link, capture_output=True, text=True)
        if res.returncode != 0:
            sys.stderr.write(res.stdout + res.stderr)
            raise RuntimeError(f"link failed with status {res.returncode}")
    os.replace(tmp, LIB_PATH)
    return LIB_PATH


if __name__ == "__main__":
    build()
