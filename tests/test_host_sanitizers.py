"""The library's host code (setup.cpp, geo.cpp) built with g++ under
AddressSanitizer and UndefinedBehaviorSanitizer and run on the CPU
(tests/host_sanitize_check.cpp): mesh, DOF, CSR, transfer, colour and slot
maps for Q1/Q2 on one and several macro cells, vertex-lattice refinement,
and the GEO reader on good and malformed files.  No GPU code is compiled."""
import os
import shutil
import subprocess

import pytest

from tests.conftest import REPO

CSRC = os.path.join(REPO, "parmoon_cmg_amd", "csrc")


@pytest.mark.skipif(shutil.which("g++") is None or not os.path.isdir("/opt/rocm/include"),
                    reason="g++ or the ROCm headers are missing")
def test_host_code_is_clean_under_asan_and_ubsan(tmp_path):
    exe = str(tmp_path / "host_sanitize_check")
    flags = ["-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
             "-fno-sanitize-recover=undefined", "-static-libasan", "-static-libubsan", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include", "-I" + CSRC]
    srcs = [os.path.join(REPO, "tests", "host_sanitize_check.cpp"), os.path.join(CSRC, "setup.cpp"),
            os.path.join(CSRC, "geo.cpp")]
    p = subprocess.run(["g++"] + flags + srcs + ["-o", exe], capture_output=True, text=True, timeout=600)
    assert p.returncode == 0, p.stderr[-3000:]
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")
    out = subprocess.run([exe, str(tmp_path)], capture_output=True, text=True, timeout=600, env=env)
    print(out.stdout[-2000:], out.stderr[-2000:])
    assert out.returncode == 0 and out.stdout.strip().endswith("OK")
    assert "ERROR: AddressSanitizer" not in out.stderr and "runtime error" not in out.stderr
