"""Compile tests/hostcheck/libbssn_cpu.so: every HIP source of the library
built as ordinary C++ for the host (shim/hip/hip_runtime.h runs each
launch on host threads) with AddressSanitizer, so that every global and
LDS index of every kernel is bounds-checked.  Test infrastructure only.

The library exposes the same C ABI as jax_bssn_amd/libbssn_amd.so, with
host pointers in place of device pointers.  Load it in a process started
with the sanitizer runtime preloaded (``asan_env()``).
"""

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
CSRC = os.path.join(ROOT, "jax_bssn_amd", "csrc")
LIB = os.path.join(HERE, "libbssn_cpu.so")
SHIM = os.path.join(HERE, "shim")

sys.path.insert(0, ROOT)
from jax_bssn_amd.build import SOURCES  # noqa: E402

FLAGS = ["-std=c++20", "-O1", "-g", "-fPIC", "-fsanitize=address", "-fno-omit-frame-pointer",
         "-Wno-unknown-pragmas", "-Wno-ignored-attributes", "-Wno-attributes"]


def deps():
    out = [os.path.join(HERE, "hipcpu.cpp"), os.path.join(SHIM, "hip", "hip_runtime.h"),
           os.path.join(ROOT, "include", "bssn_amd.h")]
    out += [os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith((".hip", ".h"))]
    return out


def stale():
    if not os.path.exists(LIB):
        return True
    t = os.path.getmtime(LIB)
    return any(os.path.getmtime(d) > t for d in deps())


def build(jobs=8):
    """One object per source (in parallel), then the link."""
    import tempfile
    srcs = [os.path.join(HERE, "hipcpu.cpp")] + [os.path.join(CSRC, s) for s in SOURCES]
    with tempfile.TemporaryDirectory(prefix="bssn_cpu_") as tmp:
        procs, objs = [], []
        for src in srcs:
            obj = os.path.join(tmp, os.path.basename(src) + ".o")
            cmd = ["g++", *FLAGS, f"-I{SHIM}", f"-I{ROOT}/include", f"-I{CSRC}", "-x", "c++", "-c", src, "-o", obj]
            while len([p for p in procs if p[1].poll() is None]) >= jobs:
                procs[[i for i, p in enumerate(procs) if p[1].poll() is None][0]][1].wait()
            procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)))
            objs.append(obj)
        failed = []
        for src, p in procs:
            out, _ = p.communicate()
            if p.returncode != 0:
                failed.append((src, out))
        if failed:
            raise RuntimeError("\n".join(f"--- {s}\n{o[-4000:]}" for s, o in failed))
        subprocess.run(["g++", "-shared", "-fsanitize=address", *objs, "-o", LIB + ".tmp", "-lpthread"], check=True)
    os.replace(LIB + ".tmp", LIB)
    return LIB


def asan_env():
    """Environment of a process that loads the library: the sanitizer
    runtime preloaded, leak checks off (the interpreter's own allocations)."""
    lib = subprocess.run(["g++", "-print-file-name=libasan.so"], capture_output=True, text=True).stdout.strip()
    env = dict(os.environ)
    env["LD_PRELOAD"] = lib
    env["ASAN_OPTIONS"] = "detect_leaks=0:abort_on_error=1:halt_on_error=1"
    return env


if __name__ == "__main__":
    print(build())
