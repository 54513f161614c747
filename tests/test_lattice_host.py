"""The lattice form's index arithmetic, checked on the host: lattice.hip's
__host__ __device__ node functions are compiled for the CPU (hipcc, host code
only runs) together with setup.cpp, and compared with the CSR matrices and
transfer operators setup.cpp builds (Q1 and Q2, one and several macro
bricks).  No GPU is needed."""
import os
import subprocess

import pytest

from tests.conftest import REPO

CSRC = os.path.join(REPO, "parmoon_cmg_amd", "csrc")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc is not installed")
def test_lattice_index_math_matches_csr_and_transfers(tmp_path):
    exe = str(tmp_path / "lattice_host_check")
    subprocess.run([HIPCC, "--offload-arch=gfx950", "-O1", "-std=c++17", "-Xclang", "-target-feature", "-Xclang",
                    "-packed-fp32-ops", "-I", CSRC, os.path.join(REPO, "tests", "lattice_host_check.cpp"),
                    os.path.join(CSRC, "setup.cpp"), "-o", exe], check=True, capture_output=True, timeout=600)
    out = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    print(out.stdout)
    assert out.returncode == 0 and out.stdout.strip().endswith("OK"), out.stdout[-3000:]
