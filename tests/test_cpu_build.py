"""Every kernel family of the library run on the CPU under AddressSanitizer.

tests/hostcheck/cpu_build.py compiles all HIP sources as host C++ (a launch
runs its blocks on host threads, __syncthreads a barrier of the block) with
-fsanitize=address, and tests/hostcheck/cpu_cases.py runs each case against
the FP64 oracle in a subprocess that preloads the sanitizer runtime: an
out-of-range global or LDS index in any kernel -- tile and face-tile
stencils, region tiles, the constraint-damping prepass on tiles and on
points, the Cartoon reconstruction, the dual-number tangent kernels, the
batched persistent launch -- aborts the case with the sanitizer's report.
The poisoned-guard case writes NaN into every sample beyond the stencils'
reach of the owned points and requires the owned RHS to be bitwise
unchanged."""

import os
import subprocess
import sys

import pytest

from tests.hostcheck import cpu_build
from tests.hostcheck.cpu_cases import CASES

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def case_runs():
    """Build the library if needed, then start every case in a process of
    its own at once (they are independent); each test collects its own."""
    if cpu_build.stale():
        cpu_build.build()
    env = cpu_build.asan_env()
    return {case: subprocess.Popen([sys.executable, "-m", "tests.hostcheck.cpu_cases", case], cwd=ROOT, env=env,
                                   stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True)
            for case in sorted(CASES)}


@pytest.mark.parametrize("case", sorted(CASES))
def test_case_under_address_sanitizer(case_runs, case):
    proc = case_runs[case]
    try:
        out, err = proc.communicate(timeout=900)
    except subprocess.TimeoutExpired:
        proc.kill()
        raise
    print(out)
    assert "AddressSanitizer" not in err, err[-6000:]
    assert proc.returncode == 0, out[-3000:] + err[-3000:]
