"""GPU parity of the plain FP64 HIP sweep (calls libminisweep.so's C-ABI).

Checked against: the compiled reference's golden vectors (tests/golden/),
the CPU oracle on seeded non-affine inputs, and the manufactured problem's
exact answer (vo == vi, normsq in closed form).
"""
import glob
import os

import numpy as np
import pytest

import oracle
import minisweep_amd as msw
from parity import (TESTER_SHAPES, assert_close, closed_form_normsq, elem_err,
                    NORM_TOL)

GOLDEN = sorted(glob.glob(os.path.join(os.path.dirname(__file__), "golden", "*.npz")))

pytestmark = pytest.mark.gpu


ENGINES = [msw.ENGINE_MFMA, msw.ENGINE_VALU]
ENGINE_IDS = ["mfma", "valu"]


def _run_gpu(d, vi0, niterations, torch, device, engine=msw.ENGINE_MFMA):
    a, m = msw.init_matrices(d.nm, d.na)
    ta = torch.from_numpy(a).to(device)
    tm = torch.from_numpy(m).to(device)
    vi = torch.from_numpy(np.ascontiguousarray(vi0)).to(device)
    vo = torch.full_like(vi, float("nan"))  # the sweep must not depend on vo
    plan = msw.Plan(d, engine=engine)
    for it in range(niterations):
        if it % 2 == 0:
            plan.sweep(vi, vo, ta, tm)
        else:
            plan.sweep(vo, vi, ta, tm)
    torch.cuda.synchronize()
    n, nd = msw.state_norms(d, vi, vo)
    plan.close()
    return vi.cpu().numpy(), vo.cpu().numpy(), n, nd


@pytest.mark.parametrize("engine", ENGINES, ids=ENGINE_IDS)
@pytest.mark.parametrize("path", GOLDEN, ids=[os.path.basename(p)[:-4] for p in GOLDEN])
def test_golden_reference_vectors(path, engine, cuda):
    import torch
    g = np.load(path)
    ncx, ncy, ncz, ne, nm, na, nu, nit = (int(x) for x in g["dims"])
    print(f"case {os.path.basename(path)} engine {engine}", flush=True)
    d = msw.dims(ncx, ncy, ncz, ne, nm, na, nu)
    od = oracle.dims(ncx, ncy, ncz, ne, nm, na, nu)
    vi0 = g["vi_in"] if "vi_in" in g else oracle.init_state(od)
    vi, vo, n, nd = _run_gpu(d, vi0, nit, torch, cuda, engine)
    if "vo" in g:
        # vo holds S^nit(vi0) for odd nit, S^(nit-1) for even; vi the other
        random_input = "_p2" in os.path.basename(path)
        n_vo = nit if nit % 2 == 1 else nit - 1
        n_vi = nit - 1 if nit % 2 == 1 else nit
        assert_close(vo, g["vo"], what="vo vs reference",
                     composed=n_vo if random_input else 1)
        if nit > 1:
            assert_close(vi, g["vi"], what="vi vs reference",
                         composed=n_vi if random_input else 1)
    assert abs(n - g["normsq"]) <= NORM_TOL * g["normsq"]
    if g["normsqdiff"] == 0:
        assert nd == 0.0
    else:
        assert abs(nd - g["normsqdiff"]) <= NORM_TOL * g["normsqdiff"]


@pytest.mark.parametrize("engine", ENGINES, ids=ENGINE_IDS)
@pytest.mark.parametrize("shape", TESTER_SHAPES, ids=[str(s) for s in TESTER_SHAPES])
def test_manufactured_exact(shape, engine, cuda):
    """vo == vi bit-exactly and normsq equals the closed form."""
    import torch
    ncx, ncy, ncz, ne, na, nm = shape
    print(f"case {shape} engine {engine}", flush=True)
    d = msw.dims(ncx, ncy, ncz, ne, nm, na, 4)
    vi_t = torch.empty(msw.size_state(d), dtype=torch.float64, device=cuda)
    msw.init_state(d, vi_t)
    vi0 = vi_t.cpu().numpy()
    od = oracle.dims(ncx, ncy, ncz, ne, nm, na, 4)
    np.testing.assert_array_equal(vi0, oracle.init_state(od))
    vi, vo, n, nd = _run_gpu(d, vi0, 1, torch, cuda, engine)
    np.testing.assert_array_equal(vo, vi0)
    assert nd == 0.0
    assert n == float(closed_form_normsq(ncx, ncy, ncz, ne, nm, 4))


RANDOM_SHAPES = [
    # ncx, ncy, ncz, ne, nm, na, nu
    (1, 1, 1, 1, 4, 1, 1),
    (9, 7, 3, 2, 16, 64, 4),
    (8, 8, 8, 2, 16, 32, 4),
    (17, 3, 5, 1, 8, 40, 2),
    (3, 20, 9, 3, 4, 13, 3),
    (16, 16, 1, 1, 16, 96, 1),
    (12, 10, 11, 2, 16, 64, 4),
]


@pytest.mark.parametrize("engine", ENGINES, ids=ENGINE_IDS)
@pytest.mark.parametrize("shape", RANDOM_SHAPES, ids=[str(s) for s in RANDOM_SHAPES])
def test_random_vs_oracle(shape, engine, cuda):
    import torch
    ncx, ncy, ncz, ne, nm, na, nu = shape
    print(f"case {shape} engine {engine}", flush=True)
    d = msw.dims(*shape)
    od = oracle.dims(*shape)
    rng = np.random.default_rng(hash(shape) & 0xFFFF)
    base = oracle.init_state(od)
    vi0 = base * (1.0 + 0.25 * rng.uniform(-1, 1, base.size))
    vi, vo, n, nd = _run_gpu(d, vi0, 1, torch, cuda, engine)
    a, m = oracle.init_matrices(od)
    vo_ref = oracle.sweep(od, a, m, vi0)
    assert_close(vo, vo_ref, what="vo vs oracle")
    n_ref, nd_ref = oracle.norms(od, vi0, vo_ref)
    assert abs(n - n_ref) <= NORM_TOL * n_ref
    assert abs(nd - nd_ref) <= NORM_TOL * nd_ref


STREAM_SHAPES = [(9, 7, 3, 3, 16, 64, 4), (16, 4, 5, 4, 16, 32, 1), (5, 9, 1, 5, 4, 40, 2)]


@pytest.mark.parametrize("shape", STREAM_SHAPES, ids=[str(s) for s in STREAM_SHAPES])
def test_energy_streaming_vs_oracle(shape, cuda):
    """Energy groups streamed through one tile pipeline (1, 2, 3 and all
    groups per workgroup): FP64 within the parity bar of the oracle, and the
    same bits for every grouping in FP64 and in the mixed FP32 form."""
    import torch
    print(f"case {shape}", flush=True)
    d = msw.dims(*shape)
    od = oracle.dims(*shape)
    rng = np.random.default_rng(17)
    base = oracle.init_state(od)
    vi0 = base * (1.0 + 0.25 * rng.uniform(-1, 1, base.size))
    a, m = oracle.init_matrices(od)
    ref = oracle.sweep(od, a, m, vi0)
    ta, tm = torch.from_numpy(a).to(cuda), torch.from_numpy(m).to(cuda)
    vi = torch.from_numpy(vi0).to(cuda)
    for prec in (msw.PREC_FP64, msw.PREC_MIXED_FP32):
        outs = []
        for epw in (1, 2, 3, shape[3]):
            vo = torch.full_like(vi, float("nan"))
            plan = msw.Plan(d, prec, engine=msw.ENGINE_MFMA, groups_per_workgroup=epw)
            plan.sweep(vi, vo, ta, tm)
            torch.cuda.synchronize()
            plan.close()
            outs.append(vo.cpu().numpy())
        if prec == msw.PREC_FP64:
            assert_close(outs[0], ref, what="vo vs oracle")
        for o in outs[1:]:
            np.testing.assert_array_equal(o, outs[0])


ROWS_SHAPES = [(9, 7, 3, 2, 16, 64, 4), (12, 20, 5, 2, 16, 32, 4), (5, 9, 4, 3, 4, 40, 2),
               (20, 12, 4, 2, 16, 32, 2), (34, 18, 3, 3, 8, 16, 1)]


@pytest.mark.parametrize("shape", ROWS_SHAPES, ids=[str(s) for s in ROWS_SHAPES])
def test_tile_rows_vs_oracle(shape, cuda):
    """Tiles of 4, 8 and 16 rows (row groups handing y faces over in LDS),
    partial tiles included: FP64 within the parity bar, and the same bits for
    every tile height, FP64 and mixed FP32 (32 angles per wave, so that 8-row
    tiles run two angle chunks by two row groups)."""
    import torch
    print(f"case {shape}", flush=True)
    d = msw.dims(*shape)
    od = oracle.dims(*shape)
    rng = np.random.default_rng(23)
    base = oracle.init_state(od)
    vi0 = base * (1.0 + 0.25 * rng.uniform(-1, 1, base.size))
    a, m = oracle.init_matrices(od)
    ref = oracle.sweep(od, a, m, vi0)
    ta, tm = torch.from_numpy(a).to(cuda), torch.from_numpy(m).to(cuda)
    vi = torch.from_numpy(vi0).to(cuda)
    for prec in (msw.PREC_FP64, msw.PREC_MIXED_FP32):
        outs = []
        for rows in (4, 8, 16):
            nac = 32 if rows <= 8 else 64
            vo = torch.full_like(vi, float("nan"))
            plan = msw.Plan(d, prec, engine=msw.ENGINE_MFMA, angles_per_wave=nac, tile_rows=rows)
            plan.sweep(vi, vo, ta, tm)
            torch.cuda.synchronize()
            plan.close()
            outs.append(vo.cpu().numpy())
        if prec == msw.PREC_FP64:
            for o in outs:
                assert_close(o, ref, what="vo vs oracle")
        for o in outs[1:]:
            if prec == msw.PREC_MIXED_FP32:
                # chunking (32 vs 64 angles per wave) changes only the order of
                # the chunk sum
                assert np.max(np.abs(o - outs[0])) <= 1e-5 * np.max(np.abs(outs[0]))
            else:
                assert_close(o, outs[0], what="tile heights")


ILL_SHAPES = [(9, 13, 9, 8, 4, 32, 4), (12, 10, 11, 2, 16, 64, 4), (17, 3, 5, 1, 8, 40, 2)]


@pytest.mark.parametrize("shape", ILL_SHAPES, ids=[str(s) for s in ILL_SHAPES])
def test_ill_conditioned_input_vs_long_double(shape, cuda):
    """The manufactured state perturbed by ±2^-8: the sweep amplifies the
    perturbation ~400x through the cancelling rows of A and M, and the
    reference's FP64 formula (the oracle) lands 1e-12-level per element from
    the exact answer (the long-double evaluation).  The GPU sweep must be at
    least as close to it: per element and normwise within twice the oracle's
    own FP64 error, or 1e-12, whichever is larger."""
    import torch
    from parity import normwise_err
    print(f"case ill-conditioned {shape}", flush=True)
    d = msw.dims(*shape)
    od = oracle.dims(*shape)
    base = oracle.init_state(od)
    x = base * (1.0 + np.random.default_rng(19).uniform(-2.0 ** -8, 2.0 ** -8, base.size))
    _, got, _, _ = _run_gpu(d, x, 1, torch, cuda)
    a, m = oracle.init_matrices(od)
    ld = oracle.sweep_ld(od, a, m, x)
    f64 = oracle.sweep(od, a, m, x)
    e_gpu, e_ref = elem_err(got, ld), elem_err(f64, ld)
    n_gpu, n_ref = normwise_err(got, ld), normwise_err(f64, ld)
    print(f"  vs long double: GPU elem {e_gpu:.3e} normwise {n_gpu:.3e}; "
          f"oracle elem {e_ref:.3e} normwise {n_ref:.3e}", flush=True)
    assert e_gpu <= max(2 * e_ref, 1e-12)
    assert n_gpu <= max(2 * n_ref, 1e-14)
