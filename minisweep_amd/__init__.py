"""minisweep_amd -- MI355X-native sweep library, Python host binding.

Thin ctypes binding over ``lib/libminisweep.so`` (the C-ABI declared in
``include/msw_plan.h``).  Device buffers come from PyTorch (plumbing only:
HBM allocation and streams); all sweep arithmetic runs in the library's HIP
kernels.  There is no CPU fallback: if the library is missing every entry
point raises.
"""
from __future__ import annotations

import ctypes
import os
from typing import Optional

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libminisweep.so")

PREC_FP64 = 0
PREC_MIXED_FP32 = 1
PREC_MIXED_BF16 = 2

ENGINE_VALU = 0
ENGINE_MFMA = 1

NOCTANT = 8


class MswDims(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int) for n in
                ("ncell_x", "ncell_y", "ncell_z", "ne", "nm", "na", "nu")]

    def tuple(self):
        return (self.ncell_x, self.ncell_y, self.ncell_z, self.ne, self.nm,
                self.na, self.nu)


class MswPlanOptions(ctypes.Structure):
    _fields_ = [("precision", ctypes.c_int), ("engine", ctypes.c_int),
                ("angles_per_wave", ctypes.c_int), ("groups_per_workgroup", ctypes.c_int),
                ("tile_rows", ctypes.c_int), ("z_block", ctypes.c_int),
                ("schedule_kind", ctypes.c_int), ("energy_offset", ctypes.c_int)]


class MswBatchOptions(ctypes.Structure):
    _fields_ = [("precision", ctypes.c_int), ("use_graph", ctypes.c_int),
                ("schedule_kind", ctypes.c_int), ("stagger", ctypes.c_int),
                ("row_groups", ctypes.c_int), ("stream_target", ctypes.c_int),
                ("xcd_map", ctypes.c_int)]


class MswServerOptions(ctypes.Structure):
    _fields_ = [("precision", ctypes.c_int), ("hbm_budget_bytes", ctypes.c_size_t),
                ("max_active", ctypes.c_int), ("keep_outputs", ctypes.c_int),
                ("schedule_kind", ctypes.c_int), ("xcd_map", ctypes.c_int)]


class MswResult(ctypes.Structure):
    _fields_ = [("tag", ctypes.c_longlong), ("normsq", ctypes.c_double),
                ("normsqdiff", ctypes.c_double), ("t_submit", ctypes.c_double),
                ("t_admit", ctypes.c_double), ("t_done", ctypes.c_double),
                ("status", ctypes.c_int)]


class SweepError(RuntimeError):
    pass


_lib = None


def lib():
    """Load the native library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise SweepError(
                f"native library missing: {LIB_PATH} (run __graft_entry__.build())")
        L = ctypes.CDLL(LIB_PATH)
        vp = ctypes.c_void_p
        dims_p = ctypes.POINTER(MswDims)
        L.msw_last_error.restype = ctypes.c_char_p
        L.msw_size_state.argtypes = [dims_p]
        L.msw_size_state.restype = ctypes.c_size_t
        L.msw_flops.argtypes = [dims_p, ctypes.c_int]
        L.msw_flops.restype = ctypes.c_double
        L.msw_plan_create.argtypes = [dims_p, ctypes.c_int, ctypes.POINTER(vp)]
        L.msw_plan_create_ex.argtypes = [dims_p, ctypes.POINTER(MswPlanOptions),
                                         ctypes.POINTER(vp)]
        L.msw_plan_default_options.argtypes = [ctypes.POINTER(MswPlanOptions)]
        L.msw_plan_workspace_bytes.argtypes = [vp]
        L.msw_plan_workspace_bytes.restype = ctypes.c_size_t
        L.msw_plan_num_launches.argtypes = [vp]
        L.msw_plan_sweep.argtypes = [vp, vp, vp, vp, vp, vp]
        L.msw_plan_sweep_host.argtypes = [vp, vp, vp, vp, vp, vp]
        L.msw_plan_sweep_faces.argtypes = [vp, vp, vp, vp, vp, vp, vp, vp, vp]
        L.msw_plan_destroy.argtypes = [vp]
        L.msw_plan_set_timing.argtypes = [vp, ctypes.c_int]
        L.msw_plan_last_timings.argtypes = [vp, ctypes.POINTER(ctypes.c_float),
                                            ctypes.c_int]
        L.msw_state_norms.argtypes = [dims_p, vp, vp, vp, vp]
        L.msw_state_norms_noalloc.argtypes = [dims_p, vp, vp, vp, vp]
        L.msw_init_state.argtypes = [dims_p, vp, vp]
        L.msw_init_matrices_host.argtypes = [ctypes.c_int, ctypes.c_int,
                                             ctypes.POINTER(ctypes.c_double),
                                             ctypes.POINTER(ctypes.c_double)]
        # batched and served sweeps (msw_batch.h)
        L.msw_batch_default_options.argtypes = [ctypes.POINTER(MswBatchOptions)]
        L.msw_batch_create.argtypes = [ctypes.c_int, dims_p, ctypes.POINTER(MswBatchOptions),
                                       ctypes.POINTER(vp)]
        L.msw_batch_set_problem.argtypes = [vp, ctypes.c_int, vp, vp, vp, vp]
        L.msw_batch_run.argtypes = [vp, vp]
        L.msw_batch_autotune.argtypes = [vp, vp, ctypes.c_int]
        L.msw_batch_set_class_nac.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.msw_batch_num_kernel_launches.argtypes = [vp]
        L.msw_batch_workspace_bytes.argtypes = [vp]
        L.msw_batch_workspace_bytes.restype = ctypes.c_size_t
        L.msw_batch_describe.argtypes = [vp, ctypes.c_char_p, ctypes.c_int]
        L.msw_batch_destroy.argtypes = [vp]
        L.msw_server_default_options.argtypes = [ctypes.POINTER(MswServerOptions)]
        L.msw_server_create.argtypes = [ctypes.POINTER(MswServerOptions), ctypes.POINTER(vp)]
        L.msw_server_submit.argtypes = [vp, dims_p, ctypes.c_longlong,
                                        ctypes.POINTER(ctypes.c_double)]
        L.msw_server_poll.argtypes = [vp, ctypes.POINTER(MswResult), ctypes.c_int]
        L.msw_server_fetch.argtypes = [vp, ctypes.c_longlong, ctypes.POINTER(ctypes.c_double)]
        L.msw_server_drain.argtypes = [vp, ctypes.c_double]
        L.msw_server_tune_from_batch.argtypes = [vp, vp]
        L.msw_server_now.argtypes = [vp]
        L.msw_server_now.restype = ctypes.c_double
        L.msw_server_stats.argtypes = [vp, ctypes.POINTER(ctypes.c_double)]
        L.msw_server_destroy.argtypes = [vp]
        _lib = L
    return _lib


def _check(rc: int, what: str):
    if rc != 0:
        msg = lib().msw_last_error().decode(errors="replace")
        raise SweepError(f"{what} failed (code {rc}): {msg}")


def dims(ncell_x, ncell_y, ncell_z, ne, nm, na, nu=4) -> MswDims:
    return MswDims(ncell_x, ncell_y, ncell_z, ne, nm, na, nu)


def size_state(d: MswDims) -> int:
    return int(lib().msw_size_state(ctypes.byref(d)))


def flops(d: MswDims, niterations: int = 1) -> float:
    return float(lib().msw_flops(ctypes.byref(d), niterations))


def init_matrices(nm: int, na: int):
    """a_from_m, m_from_a as numpy float64 (Quantities_init_am_matrices_)."""
    a = np.zeros(NOCTANT * nm * na)
    m = np.zeros(NOCTANT * nm * na)
    dp = ctypes.POINTER(ctypes.c_double)
    _check(lib().msw_init_matrices_host(nm, na, a.ctypes.data_as(dp),
                                        m.ctypes.data_as(dp)),
           "msw_init_matrices_host")
    return a, m


def _stream_ptr(stream) -> Optional[int]:
    if stream is None:
        import torch
        stream = torch.cuda.current_stream()
    return int(stream.cuda_stream) if hasattr(stream, "cuda_stream") else int(stream)


def _dptr(t) -> int:
    import torch
    assert t.is_cuda and t.dtype == torch.float64 and t.is_contiguous()
    return int(t.data_ptr())


class Plan:
    """A sweep plan: schedule + face buffers for one problem shape."""

    def __init__(self, d: MswDims, precision: int = PREC_FP64,
                 engine: Optional[int] = None, angles_per_wave: int = 0,
                 groups_per_workgroup: int = 0, tile_rows: int = 0, z_block: int = 0,
                 schedule_kind: int = 0, energy_offset: int = 0):
        self.d = d
        self.precision = precision
        o = MswPlanOptions()
        lib().msw_plan_default_options(ctypes.byref(o))
        o.precision = precision
        if engine is not None:
            o.engine = engine
        o.angles_per_wave = angles_per_wave
        o.groups_per_workgroup = groups_per_workgroup
        o.tile_rows = tile_rows
        o.z_block = z_block
        o.schedule_kind = schedule_kind
        o.energy_offset = energy_offset
        self.engine = o.engine
        h = ctypes.c_void_p()
        _check(lib().msw_plan_create_ex(ctypes.byref(d), ctypes.byref(o), ctypes.byref(h)),
               "msw_plan_create_ex")
        self._h = h

    @property
    def workspace_bytes(self) -> int:
        return int(lib().msw_plan_workspace_bytes(self._h))

    @property
    def num_launches(self) -> int:
        return int(lib().msw_plan_num_launches(self._h))

    def sweep(self, vi, vo, a_from_m, m_from_a, stream=None):
        """vo <- S(vi); all tensors float64 on the current device."""
        _check(lib().msw_plan_sweep(self._h, _dptr(vi), _dptr(vo), _dptr(a_from_m),
                                    _dptr(m_from_a), _stream_ptr(stream)),
               "msw_plan_sweep")

    def sweep_faces(self, vi, vo, a_from_m, m_from_a, facexy, facexz, faceyz, stream=None):
        """sweep() that also writes the outflow faces at the downstream
        boundary in the reference's Faces layouts (msw_plan_sweep_faces)."""
        _check(lib().msw_plan_sweep_faces(self._h, _dptr(vi), _dptr(vo), _dptr(a_from_m),
                                          _dptr(m_from_a), _dptr(facexy), _dptr(facexz),
                                          _dptr(faceyz), _stream_ptr(stream)),
               "msw_plan_sweep_faces")

    def sweep_host(self, vi_host, vo_host, a_from_m, m_from_a, stream=None):
        """vo_host <- S(vi_host) with the state in host memory (CPU tensors,
        pinned for overlap), streamed through HBM by z block; the matrices are
        device tensors.  Enqueued on stream."""
        _check(lib().msw_plan_sweep_host(self._h, ctypes.c_void_p(vi_host.data_ptr()),
                                         ctypes.c_void_p(vo_host.data_ptr()), _dptr(a_from_m),
                                         _dptr(m_from_a), _stream_ptr(stream)),
               "msw_plan_sweep_host")

    def set_timing(self, enable: bool = True):
        _check(lib().msw_plan_set_timing(self._h, int(enable)), "msw_plan_set_timing")

    def last_timings_ms(self):
        """Per-launch kernel milliseconds of the last timed sweep."""
        n = self.num_launches
        buf = (ctypes.c_float * n)()
        got = lib().msw_plan_last_timings(self._h, buf, n)
        if got < 0:
            _check(-got, "msw_plan_last_timings")
        return [float(buf[i]) for i in range(got)]

    def close(self):
        if self._h:
            lib().msw_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def state_norms(d: MswDims, vi, vo, stream=None):
    """(normsq, normsqdiff) in FP64 on the device."""
    import torch
    out = torch.empty(2, dtype=torch.float64, device=vi.device)
    _check(lib().msw_state_norms(ctypes.byref(d), _dptr(vi), _dptr(vo), _dptr(out),
                                 _stream_ptr(stream)), "msw_state_norms")
    o = out.cpu()
    return float(o[0]), float(o[1])


def state_norms_noalloc(d: MswDims, vi, vo, out, stream=None):
    """(normsq, normsqdiff) written into out (2 float64 on the device) with
    no allocation in the library call (msw_state_norms_noalloc)."""
    _check(lib().msw_state_norms_noalloc(ctypes.byref(d), _dptr(vi), _dptr(vo), _dptr(out),
                                         _stream_ptr(stream)), "msw_state_norms_noalloc")


def init_state(d: MswDims, v, stream=None):
    _check(lib().msw_init_state(ctypes.byref(d), _dptr(v), _stream_ptr(stream)),
           "msw_init_state")


def state_shape(d: MswDims):
    """Shape of the reference state layout v[iz][ie][iy][ix][iu][im]."""
    return (d.ncell_z, d.ne, d.ncell_y, d.ncell_x, d.nu, d.nm)


class Batch:
    """Many independent sweeps run together (msw_batch.h): one hipGraph replay
    per run, kernel variants chosen per (nm, na) class, optionally autotuned."""

    def __init__(self, dims_list, precision: int = PREC_FP64, use_graph: bool = True,
                 schedule_kind: int = 0, stagger: bool = True, row_groups: int = 0,
                 stream_target: int = 0, xcd_map: bool = True):
        o = MswBatchOptions()
        lib().msw_batch_default_options(ctypes.byref(o))
        o.precision, o.use_graph = precision, int(use_graph)
        o.schedule_kind, o.stagger = schedule_kind, int(stagger)
        o.row_groups, o.stream_target = row_groups, stream_target
        o.xcd_map = int(xcd_map)
        arr = (MswDims * len(dims_list))(*dims_list)
        h = ctypes.c_void_p()
        _check(lib().msw_batch_create(len(dims_list), arr, ctypes.byref(o), ctypes.byref(h)),
               "msw_batch_create")
        self._h = h
        self.dims = list(dims_list)

    def set_problem(self, i, vi, vo, a_from_m, m_from_a):
        _check(lib().msw_batch_set_problem(self._h, i, _dptr(vi), _dptr(vo), _dptr(a_from_m),
                                           _dptr(m_from_a)), "msw_batch_set_problem")

    def run(self, stream=None):
        _check(lib().msw_batch_run(self._h, _stream_ptr(stream)), "msw_batch_run")

    def autotune(self, reps: int = 2, stream=None):
        _check(lib().msw_batch_autotune(self._h, _stream_ptr(stream), reps), "msw_batch_autotune")

    def set_class_nac(self, nm: int, na: int, angles_per_wave: int):
        _check(lib().msw_batch_set_class_nac(self._h, nm, na, angles_per_wave),
               "msw_batch_set_class_nac")

    def describe(self) -> dict:
        import json
        buf = ctypes.create_string_buffer(1 << 16)
        _check(lib().msw_batch_describe(self._h, buf, len(buf)), "msw_batch_describe")
        return json.loads(buf.value.decode())

    @property
    def num_kernel_launches(self) -> int:
        return int(lib().msw_batch_num_kernel_launches(self._h))

    @property
    def workspace_bytes(self) -> int:
        return int(lib().msw_batch_workspace_bytes(self._h))

    def close(self):
        if self._h:
            lib().msw_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Request loop with continuous batching (msw_batch.h): submit problems,
    poll completions (FP64 norms and server-clock timestamps)."""

    def __init__(self, precision: int = PREC_FP64, hbm_budget_bytes: int = 32 << 30,
                 max_active: int = 512, keep_outputs: bool = False, schedule_kind: int = 0,
                 xcd_map: bool = True):
        o = MswServerOptions()
        lib().msw_server_default_options(ctypes.byref(o))
        o.precision, o.hbm_budget_bytes = precision, hbm_budget_bytes
        o.max_active, o.keep_outputs = max_active, int(keep_outputs)
        o.schedule_kind, o.xcd_map = schedule_kind, int(xcd_map)
        h = ctypes.c_void_p()
        _check(lib().msw_server_create(ctypes.byref(o), ctypes.byref(h)), "msw_server_create")
        self._h = h

    def submit(self, d: MswDims, tag: int, vi_host=None):
        p = None
        if vi_host is not None:
            vi_host = np.ascontiguousarray(vi_host, dtype=np.float64)
            assert vi_host.size == size_state(d)
            p = vi_host.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        _check(lib().msw_server_submit(self._h, ctypes.byref(d), tag, p), "msw_server_submit")

    def poll(self, max_results: int = 4096):
        buf = (MswResult * max_results)()
        n = lib().msw_server_poll(self._h, buf, max_results)
        if n < 0:
            _check(-n, "msw_server_poll")
        return [{"tag": r.tag, "normsq": r.normsq, "normsqdiff": r.normsqdiff,
                 "t_submit": r.t_submit, "t_admit": r.t_admit, "t_done": r.t_done,
                 "status": r.status} for r in buf[:n]]

    def fetch(self, tag: int, d: MswDims):
        out = np.empty(size_state(d))
        _check(lib().msw_server_fetch(self._h, tag, out.ctypes.data_as(
            ctypes.POINTER(ctypes.c_double))), "msw_server_fetch")
        return out

    def drain(self, timeout_s: float = 0.0):
        _check(lib().msw_server_drain(self._h, timeout_s), "msw_server_drain")

    def tune_from(self, batch: "Batch"):
        """Serve with an autotuned batch's kernel configuration."""
        _check(lib().msw_server_tune_from_batch(self._h, batch._h), "msw_server_tune_from_batch")

    def now(self) -> float:
        return float(lib().msw_server_now(self._h))

    def stats(self) -> dict:
        out = (ctypes.c_double * 4)()
        _check(lib().msw_server_stats(self._h, out), "msw_server_stats")
        return {"done": int(out[0]), "kernel_launches": int(out[1]), "ticks": int(out[2]),
                "arena_peak_bytes": int(out[3])}

    def close(self):
        if self._h:
            lib().msw_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
