"""The batched form's host code (batch.hip: schedule, tick assembly, segment
tables, graph capture and release, the autotuner) built for the CPU with
AddressSanitizer and UndefinedBehaviorSanitizer against a stand-in HIP
runtime (tests/host_harness/fake_hip.cpp).  The stand-in keeps device memory
on the host heap and, instead of running a kernel, checks what it would
touch: every workgroup's segment lookup, problem parameters, tile item and
the extents of the buffers named in its parameters.  Timings are random, so
the autotuner walks different candidate sequences on every iteration."""
import os
import shutil
import subprocess

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(REPO, "minisweep_amd", "csrc")
HARNESS = os.path.join(REPO, "tests", "host_harness")
CXX = "/opt/rocm/llvm/bin/clang++"


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    if not os.path.exists(CXX) or not os.path.isdir("/opt/rocm/include/hip"):
        pytest.skip("clang++ or the HIP headers are not installed")
    out = str(tmp_path_factory.mktemp("harness") / "batch_host")
    cmd = [CXX, "-std=c++17", "-g", "-O1", "-fsanitize=address,undefined",
           "-fno-sanitize-recover=undefined", "-fno-omit-frame-pointer", "-Wno-unused-value",
           "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include", "-I" + os.path.join(REPO, "include"),
           "-I" + CSRC, "-x", "c++", os.path.join(CSRC, "batch.hip"),
           os.path.join(CSRC, "plan_host.cpp"), os.path.join(HARNESS, "fake_hip.cpp"),
           os.path.join(HARNESS, "driver.cpp"), "-o", out, "-lpthread"]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    return out


def test_batch_host_code_under_asan(harness):
    env = dict(os.environ, MSW_BATCH_AUTOTUNE="1",
               ASAN_OPTIONS="detect_leaks=1:abort_on_error=1")
    r = subprocess.run([harness, "1"], capture_output=True, text=True, timeout=600, env=env)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-4000:]
    assert "fails 0" in r.stdout
