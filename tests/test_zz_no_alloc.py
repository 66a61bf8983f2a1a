"""msw_state_norms_noalloc: the norms with their block partials in the
library's static device scratch (no allocation in the call), against
msw_state_norms (which allocates its partials on a thread's first call).

Both run the same kernels in the same order, so the two results are the same
bits: checked on two calls in a row, the second with a smaller input, and on
a third call after a different state, so that a stale partial left in the
shared scratch by an earlier call would show.
"""
import ctypes

import numpy as np
import pytest

import minisweep_amd as msw

pytestmark = pytest.mark.gpu


def _noalloc(d, vi, vo, out):
    msw.state_norms_noalloc(d, vi, vo, out)


def _state(torch, dev, d, seed):
    g = torch.Generator(device=dev)
    g.manual_seed(seed)
    n = msw.size_state(d)
    return (torch.rand(n, generator=g, dtype=torch.float64, device=dev),
            torch.rand(n, generator=g, dtype=torch.float64, device=dev))


def test_noalloc_norms_match_allocating_entry_point(cuda):
    import torch
    out = torch.zeros(2, dtype=torch.float64, device=cuda)
    for step, shape in enumerate([(40, 33, 21, 8, 16, 32, 4), (5, 3, 2, 1, 4, 8, 1),
                                  (17, 9, 11, 3, 8, 16, 2)]):
        print(f"case no-alloc norms call {step} shape {shape}", flush=True)
        d = msw.dims(*shape)
        vi, vo = _state(torch, cuda, d, 100 + step)
        ref = msw.state_norms(d, vi, vo)
        out.fill_(float("nan"))
        _noalloc(d, vi, vo, out)
        torch.cuda.synchronize()
        got = tuple(out.cpu().numpy().tolist())
        assert got == ref, (step, got, ref)
        # against a host sum too (order differs: FP64 rounding only)
        a, b = vi.cpu().numpy(), vo.cpu().numpy()
        np.testing.assert_allclose(got, (float(np.sum(b * b)), float(np.sum((a - b) ** 2))),
                                   rtol=1e-12)


def test_noalloc_norms_rejects_bad_arguments(cuda):
    import torch
    L = msw.lib()
    d = msw.dims(0, 1, 1, 1, 4, 8, 1)
    out = torch.zeros(2, dtype=torch.float64, device=cuda)
    rc = L.msw_state_norms_noalloc(ctypes.byref(d), ctypes.c_void_p(out.data_ptr()),
                                   ctypes.c_void_p(out.data_ptr()),
                                   ctypes.c_void_p(out.data_ptr()), None)
    assert rc != 0
