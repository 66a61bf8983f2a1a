"""The drop-in boundary as a C maintainer uses it: the reference's own driver
sources (src/4_driver/sweep.c, runner.c, tester.c), unchanged, compiled
against this repository's include/ and linked with libminisweep.so.

Runs only where the reference tree is present (it never travels with the
repository).  No GPU is needed: the linked `sweep` is run on a configuration
whose Insist fires before any device call (runner.c:89-95), and the exit
status and message must be the reference's (env_assert.c:27-32).
"""
import os
import subprocess

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_DRIVER = "/root/reference/src/4_driver"
LIBDIR = os.path.join(REPO, "minisweep_amd", "lib")
INCLUDE = os.path.join(REPO, "include")

pytestmark = pytest.mark.skipif(
    not os.path.isdir(REF_DRIVER) or not os.path.exists(os.path.join(LIBDIR, "libminisweep.so")),
    reason="needs the reference sources and the built library")

CFLAGS = ["-std=c99", "-O2", "-I" + INCLUDE, "-I/opt/rocm/include", "-D__HIP_PLATFORM_AMD__"]
LDFLAGS = ["-L" + LIBDIR, "-lminisweep", "-Wl,-rpath," + LIBDIR, "-L/opt/rocm/lib",
           "-lamdhip64", "-Wl,-rpath,/opt/rocm/lib"]


@pytest.fixture(scope="module")
def drivers(tmp_path_factory):
    out = tmp_path_factory.mktemp("relink")
    objs = {}
    for name in ("sweep", "runner", "tester"):
        o = str(out / f"{name}.o")
        r = subprocess.run(["gcc", *CFLAGS, "-c", os.path.join(REF_DRIVER, f"{name}.c"),
                            "-o", o], capture_output=True, text=True)
        assert r.returncode == 0, r.stderr
        objs[name] = o
    exes = {}
    for main in ("sweep", "tester"):
        exe = str(out / main)
        r = subprocess.run(["gcc", "-o", exe, objs[main], objs["runner"], *LDFLAGS],
                           capture_output=True, text=True)
        assert r.returncode == 0, r.stderr
        exes[main] = exe
    return exes


def test_reference_drivers_compile_and_link(drivers):
    assert os.access(drivers["sweep"], os.X_OK) and os.access(drivers["tester"], os.X_OK)


@pytest.mark.parametrize("flag,cond", [
    ("--ncell_x", 'dims_g.ncell_x > 0 ? "Invalid ncell_x supplied." : 0'),
    ("--ne", 'dims_g.ne > 0 ? "Invalid ne supplied." : 0'),
    ("--na", 'dims_g.na > 0 ? "Invalid na supplied." : 0')])
def test_relinked_sweep_insist(drivers, flag, cond):
    """The Insist macro prints its condition text (env_assert_kernels.h:30-31)
    through this library's insist_, which exits with EXIT_FAILURE."""
    r = subprocess.run([drivers["sweep"], flag, "0"], capture_output=True, text=True,
                       timeout=60)
    assert r.returncode == 1
    assert r.stderr.startswith(f'Insist error: "{cond}". At file ')
    assert "src/4_driver/runner.c, line " in r.stderr
