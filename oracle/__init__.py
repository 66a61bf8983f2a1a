"""CPU oracle for the minisweep sweep -- TEST INFRASTRUCTURE ONLY.

Loads ``oracle/build/liboracle.so`` (built by ``oracle/Makefile`` from
``oracle.c``, a C restatement of the reference algorithm; see ``oracle.h`` for
the reference file:line each function follows).  Only ``tests/``,
``__graft_entry__.smoke()`` and ``bench.py``'s cpu_baseline leg import this
package, as the checker.  The product library ``minisweep_amd`` never does.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "build", "liboracle.so")
_lib = None


class OracleDims(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int) for n in
                ("ncell_x", "ncell_y", "ncell_z", "ne", "nm", "na", "nu")]


def build() -> None:
    subprocess.check_call(["make", "-s", "-C", _HERE, "liboracle"])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        dp = ctypes.POINTER(ctypes.c_double)
        L.oracle_size_state.argtypes = [OracleDims]
        L.oracle_size_state.restype = ctypes.c_size_t
        L.oracle_init_matrices.argtypes = [OracleDims, dp, dp]
        L.oracle_init_state.argtypes = [OracleDims, dp]
        L.oracle_sweep.argtypes = [OracleDims, dp, dp, dp, dp, ctypes.c_int]
        L.oracle_sweep_tileoctants.argtypes = [OracleDims, dp, dp, dp, dp, ctypes.c_int]
        L.oracle_sweep_ld.argtypes = [OracleDims, dp, dp, dp, dp, ctypes.c_int]
        L.oracle_sweep_faces.argtypes = [OracleDims, dp, dp, dp, dp, dp, dp, dp, ctypes.c_int]
        L.oracle_sweep_range.argtypes = [OracleDims, dp, dp, dp, dp,
                                         ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.oracle_norms.argtypes = [OracleDims, dp, dp, dp, dp]
        L.oracle_face_value.argtypes = [OracleDims] + [ctypes.c_int] * 7
        L.oracle_face_value.restype = ctypes.c_double
        L.oracle_flops.argtypes = [OracleDims, ctypes.c_int]
        L.oracle_flops.restype = ctypes.c_double
        _lib = L
    return _lib


def _p(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def dims(ncell_x, ncell_y, ncell_z, ne, nm, na, nu=4) -> OracleDims:
    return OracleDims(ncell_x, ncell_y, ncell_z, ne, nm, na, nu)


def state_shape(d: OracleDims):
    """numpy shape of the state layout v[iz][ie][iy][ix][iu][im]."""
    return (d.ncell_z, d.ne, d.ncell_y, d.ncell_x, d.nu, d.nm)


def init_matrices(d: OracleDims):
    a = np.zeros(8 * d.nm * d.na)
    m = np.zeros(8 * d.nm * d.na)
    lib().oracle_init_matrices(d, _p(a), _p(m))
    return a, m


def init_state(d: OracleDims) -> np.ndarray:
    v = np.zeros(lib().oracle_size_state(d))
    lib().oracle_init_state(d, _p(v))
    return v


def sweep(d: OracleDims, a, m, vi, nthreads: int = 0) -> np.ndarray:
    vi = np.ascontiguousarray(vi, dtype=np.float64)
    vo = np.zeros_like(vi)
    lib().oracle_sweep(d, _p(np.ascontiguousarray(a)), _p(np.ascontiguousarray(m)),
                       _p(vi), _p(vo), nthreads)
    return vo


def sweep_tileoctants(d: OracleDims, a, m, vi, nthreads: int = 0) -> np.ndarray:
    """The sweep in the reference's tiled-octant order (8 disjoint
    half-domain boxes per step); equal to sweep() up to rounding."""
    vi = np.ascontiguousarray(vi, dtype=np.float64)
    vo = np.zeros_like(vi)
    lib().oracle_sweep_tileoctants(d, _p(np.ascontiguousarray(a)), _p(np.ascontiguousarray(m)),
                                   _p(vi), _p(vo), nthreads)
    return vo


def sweep_ld(d: OracleDims, a, m, vi, nthreads: int = 0) -> np.ndarray:
    """The sweep evaluated in long double, rounded to double at the end: the
    near-exact answer against which FP64 evaluations are judged on
    ill-conditioned inputs."""
    vi = np.ascontiguousarray(vi, dtype=np.float64)
    vo = np.zeros_like(vi)
    lib().oracle_sweep_ld(d, _p(np.ascontiguousarray(a)), _p(np.ascontiguousarray(m)),
                          _p(vi), _p(vo), nthreads)
    return vo


def sweep_faces(d: OracleDims, a, m, vi, nthreads: int = 0):
    """(vo, facexy, facexz, faceyz): the sweep and each octant's outflow faces
    at the downstream boundary, reference layouts with all 8 octants
    ([8][ne][ny][nx][nu][na], [8][ne][nz][nx][nu][na], [8][ne][nz][ny][nu][na]),
    scaled by (1+octant) as the reference stores them."""
    vo = np.empty_like(vi)
    nx, ny, nz, ne, na, nu = d.ncell_x, d.ncell_y, d.ncell_z, d.ne, d.na, d.nu
    fxy = np.empty(8 * ne * ny * nx * nu * na)
    fxz = np.empty(8 * ne * nz * nx * nu * na)
    fyz = np.empty(8 * ne * nz * ny * nu * na)
    lib().oracle_sweep_faces(d, _p(a), _p(m), _p(np.ascontiguousarray(vi)), _p(vo), _p(fxy),
                             _p(fxz), _p(fyz), nthreads)
    return vo, fxy, fxz, fyz


def sweep_range(d: OracleDims, a, m, vi, vo, ie_begin, ie_end, nthreads=0):
    lib().oracle_sweep_range(d, _p(a), _p(m), _p(vi), _p(vo), ie_begin, ie_end,
                             nthreads)


def run(d: OracleDims, a, m, vi, niterations: int = 1, nthreads: int = 0):
    """Ping-pong iterations as Runner_run_case (src/4_driver/runner.c:148-155).

    Returns the final (vi, vo) buffers."""
    vi = np.array(vi, dtype=np.float64, copy=True)
    vo = np.zeros_like(vi)
    for it in range(niterations):
        if it % 2 == 0:
            vo = sweep(d, a, m, vi, nthreads)
        else:
            vi = sweep(d, a, m, vo, nthreads)
    return vi, vo


def norms(d: OracleDims, vi, vo):
    n = ctypes.c_double()
    nd = ctypes.c_double()
    lib().oracle_norms(d, _p(np.ascontiguousarray(vi)), _p(np.ascontiguousarray(vo)),
                       ctypes.byref(n), ctypes.byref(nd))
    return n.value, nd.value


def face_value(d, ix, iy, iz, ie, ia, iu, octant) -> float:
    return lib().oracle_face_value(d, ix, iy, iz, ie, ia, iu, octant)


def flops(d: OracleDims, niterations: int = 1) -> float:
    return lib().oracle_flops(d, niterations)
