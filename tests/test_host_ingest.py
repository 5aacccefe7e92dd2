"""Host ingest helpers of the library (no device calls): the thread pool that
splits passes over the inputs, the fused antenna-state scan that decides the
per-chirp paths, the grid's bounding box, and the bound that lets the
range-profile tile advance ranges by FP32 increments (checked against the
same FP32 operations emulated on the host over a 1024-chirp arc). A small C++
harness links the built library and checks them against serial passes
(tests/host/ingest_check.cpp)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "sarlab-tdc_amd", "sarlabtdc")
LIB = os.path.join(PKG, "libsarlabtdc_amd.so")
HIPCC = "/opt/rocm/bin/hipcc"


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    if not os.path.exists(LIB):
        pytest.fail("libsarlabtdc_amd.so is not built (run the package build first)")
    if not os.path.exists(HIPCC) or shutil.which("g++") is None:
        pytest.skip("no host compiler")
    exe = str(tmp_path_factory.mktemp("ingest") / "ingest_check")
    cmd = [HIPCC, "-O2", "-std=c++17", "-x", "c++", "-D__HIP_PLATFORM_AMD__",
           "-I" + os.path.join(ROOT, "include"), "-I" + os.path.join(ROOT, "sarlab-tdc_amd", "csrc"),
           "-I/opt/rocm/include", os.path.join(ROOT, "tests", "host", "ingest_check.cpp"),
           "-L" + PKG, "-lsarlabtdc_amd", "-Wl,-rpath," + PKG, "-pthread", "-o", exe]
    subprocess.run(cmd, check=True, capture_output=True, text=True, timeout=300)
    return exe


def test_pool_and_antenna_scan(harness):
    r = subprocess.run([harness], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0 and r.stdout.strip().endswith("OK"), r.stdout
