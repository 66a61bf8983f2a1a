"""GPU tests of the drop-in path: the `sweep` and `tester` executables, built
on Runner_run_case -> Sweeper_create/Sweeper_sweep/Sweeper_destroy
(include/sweeper.h), i.e. the reference's own driver flow.
"""
import os
import re
import subprocess

import pytest

from parity import closed_form_normsq

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIBDIR = os.path.join(REPO, "minisweep_amd", "lib")

pytestmark = pytest.mark.gpu

RESULT = re.compile(r"Normsq result: (\S+)  diff: (\S+)  (PASS|FAIL)  time: (\S+)  GF/s: (\S+)")


def run(exe, *args, timeout=100):
    print("run", exe, *args, flush=True)
    return subprocess.run([os.path.join(LIBDIR, exe), *map(str, args)],
                          capture_output=True, text=True, timeout=timeout)


def test_sweep_no_args_banner(cuda):
    """`sweep` with no arguments: defaults 5^3 ne30 na33 (runner.c:80-85) and
    the TESTS banner (sweep.c:59-65)."""
    r = run("sweep")
    assert r.returncode == 0, r.stderr
    m = RESULT.search(r.stdout)
    assert m and m.group(3) == "PASS"
    assert float(m.group(1)) == float(f"{closed_form_normsq(5, 5, 5, 30, 16):.8e}")
    assert "TESTS 1    PASSED 1    FAILED 0" in r.stdout


@pytest.mark.parametrize("args,shape", [
    (["--ncell_x", 8, "--ncell_y", 8, "--ncell_z", 8, "--ne", 4, "--na", 32],
     (8, 8, 8, 4, 16)),
    (["--ncell_x", 5, "--ncell_y", 5, "--ncell_z", 5, "--ne", 30, "--na", 33, "--nm", 4],
     (5, 5, 5, 30, 4)),
    (["--ncell_x", 9, "--ncell_y", 4, "--ncell_z", 6, "--ne", 3, "--na", 64,
      "--niterations", 3, "--nblock_z", 3, "--nthread_octant", 8, "--nthread_e", 4],
     (9, 4, 6, 3, 16)),
])
def test_sweep_result_line(args, shape, cuda):
    r = run("sweep", *args)
    assert r.returncode == 0, r.stderr
    m = RESULT.search(r.stdout)
    assert m, r.stdout
    assert m.group(3) == "PASS" and float(m.group(2)) == 0.0
    assert m.group(1) == f"{closed_form_normsq(*shape):.8e}"


@pytest.mark.parametrize("args,msg", [
    (["--ncell_z", 5, "--nblock_z", 2], "Currently require all blocks have same z dimension"),
    (["--nthread_octant", 3], "Invalid thread count supplied"),
    (["--nsemiblock", 3], "Invalid semiblock count supplied"),
    (["--bogus", 1], "Invalid argument detected."),
    (["--nm", 5], "nm must be 4, 8 or 16"),
    (["--ncell_z", 4, "--nblock_z", 2, "--precision", 1],
     "z blocking (--nblock_z > 1) is supported in FP64 only"),
])
def test_sweep_config_errors(args, msg, cuda):
    r = run("sweep", *args)
    assert r.returncode == 1
    assert "Insist error" in r.stderr and msg in r.stderr


# pairs per battery: serial 54, openmp 981 (tester.c:103-211), tasks 64,
# cuda 984 (tester.c:290-390), mpi 4 one-process pairs, variants 2
@pytest.mark.parametrize("battery", ["serial", "openmp", "tasks", "cuda", "mpi", "variants"])
def test_tester_battery(battery, cuda):
    """The reference's compare_runs batteries (tester.c): every pair must
    give normsqdiff == 0 and identical normsq, exactly."""
    r = run("tester", "--battery", battery, timeout=110)
    print(r.stdout.strip().splitlines()[-1] if r.stdout.strip() else "(no output)", flush=True)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    m = re.search(r"TESTS (\d+)    PASSED (\d+)    FAILED (\d+)", r.stdout)
    assert m and int(m.group(1)) > 0 and m.group(1) == m.group(2) and m.group(3) == "0"
