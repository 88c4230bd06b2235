"""Host code under AddressSanitizer + UndefinedBehaviorSanitizer (CPU only).

tests/host_sanitize.cpp drives the seeded generator
(matrix_format_performance_amd/csrc/generate.cpp), the MatrixMarket reader
of include/mfp_compat.hpp and the oracle's C restatement over small inputs
and malformed files; ROCm's clang++ builds it for the host with both
sanitizers, any report aborting the run.  The GPU kernels are covered by the
parity tests instead (device sanitizers are not part of this stack).
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLANG = "/opt/rocm/llvm/bin/clang++"


@pytest.mark.skipif(not os.path.exists(CLANG) or shutil.which("true") is None, reason="ROCm clang++ not present")
def test_host_code_is_clean_under_asan_and_ubsan(tmp_path):
    exe = str(tmp_path / "host_sanitize")
    san = ["-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined", "-fno-sanitize-recover=all"]
    obj = str(tmp_path / "oracle.o")
    steps = [
        [CLANG, "-x", "c", "-std=c11", *san, "-c", os.path.join(ROOT, "oracle", "mfp_oracle.c"), "-o", obj],
        [CLANG, "-std=c++17", *san, "-D__HIP_PLATFORM_AMD__=1", "-I/opt/rocm/include",
         "-I" + os.path.join(ROOT, "include"),
         "-x", "c++", os.path.join(ROOT, "matrix_format_performance_amd", "csrc", "generate.cpp"),
         os.path.join(ROOT, "tests", "host_sanitize.cpp"), "-x", "none", obj, "-lpthread", "-lm", "-o", exe],
    ]
    for cmd in steps:
        r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
        assert r.returncode == 0, r.stderr[-3000:]
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")
    run = subprocess.run([exe], capture_output=True, text=True, timeout=300, env=env)
    assert run.returncode == 0, run.stdout[-2000:] + run.stderr[-4000:]
    assert "ALL OK" in run.stdout
    assert "runtime error" not in run.stderr and "AddressSanitizer" not in run.stderr
