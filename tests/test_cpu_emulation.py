"""The library's kernel sources run on the CPU HIP emulation under
AddressSanitizer (tests/cpu_emu/), against the FP64 oracle: every device
index is bounds-checked on the host.  No GPU needed.

Builds tests/cpu_emu/_build/libhsba_cpu.so (incremental), then runs
tests/cpu_emu/run_cases.py in a child process with the ASan runtime
preloaded.  Skipped when the ROCm clang that compiles the emulation is absent.
"""
import os
import subprocess
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
EMU = os.path.join(HERE, "cpu_emu")
CLANG = "/opt/rocm/llvm/bin/clang++"

# the tiled factorization and both solves, the matrix-core pair kernel and
# the one-workgroup factorization, batched windows, the allocation-free
# covariance (about 5 minutes in all; run_cases.py also holds x_covariance,
# the full dense covariance's Jacobi and Schur routes, ~6 minutes more, and
# fp64_wide)
CASES = ["tiled_cholesky", "small_mma", "batch", "covariance_static"]


@pytest.mark.skipif(not os.path.exists(CLANG), reason="ROCm clang++ not installed")
def test_kernels_under_address_sanitizer():
    b = subprocess.run(["bash", os.path.join(EMU, "build.sh")], capture_output=True, text=True, timeout=900)
    assert b.returncode == 0, b.stderr[-3000:]
    lib = b.stdout.strip().splitlines()[-1]
    asan = subprocess.run([CLANG, "-print-file-name=libclang_rt.asan-x86_64.so"], capture_output=True,
                          text=True).stdout.strip()
    env = dict(os.environ, LD_PRELOAD=asan, HSBA_LIB=lib, ASAN_OPTIONS="detect_leaks=0")
    r = subprocess.run([sys.executable, os.path.join(EMU, "run_cases.py"), *CASES], capture_output=True,
                       text=True, env=env, timeout=900)
    print(r.stdout)
    assert r.returncode == 0 and "EMU_OK" in r.stdout, (r.stdout[-2000:], r.stderr[-3000:])
