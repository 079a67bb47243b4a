"""The library's host code under AddressSanitizer and UndefinedBehaviorSanitizer:
tests/sanitize/host_check.cpp and the host sources (host_graph.cpp,
drivers.cpp) built with g++ (CPU code only; no device code is compiled) and
run; any leak, out-of-bounds access, mismatched free or undefined behaviour
aborts it."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(ROOT, "onegraph_amd", "csrc")
ROCM = "/opt/rocm"


def test_host_code_under_asan_and_ubsan(tmp_path):
    if not shutil.which("g++") or not os.path.exists(os.path.join(ROCM, "include", "hip", "hip_runtime.h")):
        pytest.skip("g++ or the HIP headers are missing")
    exe = str(tmp_path / "host_check")
    flags = ["-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
             "-fno-sanitize-recover=all", "-D__HIP_PLATFORM_AMD__", "-I" + os.path.join(ROCM, "include"),
             "-I" + os.path.join(ROOT, "include"), "-I" + CSRC]
    srcs = [os.path.join(HERE, "sanitize", "host_check.cpp"), os.path.join(CSRC, "host_graph.cpp"),
            os.path.join(CSRC, "drivers.cpp")]
    r = subprocess.run(["g++", *flags, *srcs, "-o", exe, "-L" + os.path.join(ROCM, "lib"), "-lamdhip64",
                        "-Wl,-rpath," + os.path.join(ROCM, "lib"), "-lpthread"], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-4000:]
    supp = tmp_path / "lsan.supp"
    supp.write_text("leak:libamdhip64\nleak:libhsa-runtime64\nleak:libhsakmt\nleak:libdrm\n")
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1",
               LSAN_OPTIONS=f"suppressions={supp}", UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    r = subprocess.run([exe, str(tmp_path)], capture_output=True, text=True, env=env, timeout=300)
    assert r.returncode == 0, (r.stdout + r.stderr)[-6000:]
    assert "host checks: 0 failures" in r.stdout
