"""The C-ABI library loads and exports every function include/*.h declares,
and the host-only parts of the API behave as the reference's (no GPU calls).
"""
import ctypes
import os
import re
import subprocess

import pytest

import minisweep_amd as msw

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INC = os.path.join(REPO, "include")
LIBDIR = os.path.join(REPO, "minisweep_amd", "lib")


def declared_functions():
    """Non-inline function declarations in include/*.h."""
    names = set()
    for fn in sorted(os.listdir(INC)):
        if not fn.endswith(".h"):
            continue
        text = open(os.path.join(INC, fn)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        text = re.sub(r"//[^\n]*", "", text)
        # drop static inline bodies
        text = re.sub(r"static inline[^{;]*\{[^}]*\}", "", text)
        for m in re.finditer(r"^[A-Za-z_][\w\s\*]*?\b([A-Za-z_]\w*)\s*\([^;{]*\)\s*;",
                             text, flags=re.M):
            name = m.group(1)
            if name not in ("if", "while", "return", "sizeof"):
                names.add(name)
    return sorted(names)


def test_headers_declare_entry_points():
    names = declared_functions()
    for must in ("Sweeper_create", "Sweeper_sweep", "Sweeper_destroy", "Sweeper_null",
                 "msw_plan_create", "msw_plan_sweep", "msw_plan_destroy",
                 "Arguments_consume_int_or_default", "Runner_run_case",
                 "compare_runs", "Quantities_create", "get_state_norms"):
        assert must in names, must


@pytest.mark.parametrize("name", declared_functions())
def test_library_exports(name):
    lib = msw.lib()
    assert hasattr(lib, name), f"{name} declared in include/ but not exported"


def _args_from_string(lib, s):
    class Arguments(ctypes.Structure):
        _fields_ = [("argc", ctypes.c_int), ("argv_unconsumed", ctypes.c_void_p),
                    ("argstring", ctypes.c_char_p)]
    a = Arguments()
    lib.Arguments_create_from_string(ctypes.byref(a), s.encode())
    return a


def test_arguments_read_once_semantics():
    """Arguments (src/1_base/arguments.c:37-198): consume once, default
    otherwise, leftovers detected."""
    lib = msw.lib()
    lib.Arguments_consume_int_or_default.restype = ctypes.c_int
    lib.Arguments_are_all_consumed.restype = ctypes.c_int
    lib.Arguments_exists.restype = ctypes.c_int
    a = _args_from_string(lib, "--ncell_x 7  --ne\t3 --bogus 1")
    assert a.argc == 7
    assert lib.Arguments_exists(ctypes.byref(a), b"--ne") == 1
    assert lib.Arguments_consume_int_or_default(ctypes.byref(a), b"--ncell_x", 5) == 7
    assert lib.Arguments_consume_int_or_default(ctypes.byref(a), b"--ncell_x", 5) == 5
    assert lib.Arguments_consume_int_or_default(ctypes.byref(a), b"--ne", 30) == 3
    assert lib.Arguments_are_all_consumed(ctypes.byref(a)) == 0
    assert lib.Arguments_consume_int_or_default(ctypes.byref(a), b"--bogus", 0) == 1
    assert lib.Arguments_are_all_consumed(ctypes.byref(a)) == 1
    lib.Arguments_destroy(ctypes.byref(a))


def test_flops_and_sizes():
    d = msw.dims(64, 64, 64, 64, 16, 64, 4)
    assert msw.size_state(d) == 2 ** 30
    assert msw.flops(d, 1) == 2611340115968.0
    assert msw.flops(d, 4) == 4 * 2611340115968.0


def test_plan_rejects_bad_dims_without_gpu():
    """Argument errors are caught before any device work, with the message
    text exact (nothing appended)."""
    with pytest.raises(msw.SweepError, match=r"all dimensions must be positive$"):
        msw.Plan(msw.dims(0, 4, 4, 1, 16, 8, 4))
    with pytest.raises(msw.SweepError, match=r"nm must be 4, 8 or 16 \(got 5\)$"):
        msw.Plan(msw.dims(4, 4, 4, 1, 5, 8, 4))
    with pytest.raises(msw.SweepError, match=r"unknown precision 7$"):
        msw.Plan(msw.dims(4, 4, 4, 1, 16, 8, 4), precision=7)


def _sweep_cli(*args):
    exe = os.path.join(LIBDIR, "sweep")
    return subprocess.run([exe, *args], capture_output=True, text=True, timeout=60)


@pytest.mark.parametrize("args,msg", [
    (["--ncell_x", "0"], "Invalid ncell_x supplied."),
    (["--ncell_y", "-1"], "Invalid ncell_y supplied."),
    (["--ne", "0"], "Invalid ne supplied."),
    (["--na", "0"], "Invalid na supplied."),
    (["--niterations", "-1"], "Invalid iteration count supplied."),
    (["--ncell_z"], "i < args->argc"),
    (["--nproc_x", "2"], "nproc_x and nproc_y must be 1"),
    (["--is_using_device", "2"], "Invalid is_using_device supplied."),
    (["--precision", "9"], "Invalid precision supplied."),
])
def test_cli_insist_errors(args, msg):
    """Configuration errors print an Insist line and exit with failure
    before any device work (src/1_base/env_assert.c:27-32)."""
    r = _sweep_cli(*args)
    assert r.returncode == 1
    assert "Insist error" in r.stderr and msg in r.stderr
