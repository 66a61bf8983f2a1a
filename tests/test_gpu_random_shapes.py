"""GPU parity over seeded random shapes and plan options, through the C-ABI.

FP64: a perturbed input (vi * (1 + U(-1/4, 1/4))) swept by a plan with
random angles per wave, tile rows and energy groups per workgroup, and by a
z-blocked plan where ncell_z has a proper divisor, must match the CPU oracle
within the FP64 parity bar.  Mixed FP32 / BF16: on the
manufactured input every operand is exact in both forms, so vo == vi and
normsq equals its closed form exactly, for any shape.

Shapes are drawn from a fixed seed: ncell 1..14 per axis, ne 1..3, nm 4 / 8 /
16, na 1..96, nu 1..4 (partial tiles in x and y, padded angle blocks,
single-cell axes).
"""
import numpy as np
import pytest

import oracle
import minisweep_amd as msw
from parity import assert_close, closed_form_normsq

pytestmark = pytest.mark.gpu


def _shapes(n, seed):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        ncx, ncy, ncz = (int(v) for v in rng.integers(1, 15, 3))
        out.append((ncx, ncy, ncz, int(rng.integers(1, 4)), int(rng.choice([4, 8, 16])),
                    int(rng.integers(1, 97)), int(rng.integers(1, 5))))
    return out


def _options(shape, rng):
    """Random angles per wave / tile rows / groups per workgroup that the
    plan builds for this shape (at most 4 angle waves, at most 4 waves per
    workgroup)."""
    na = shape[5]
    while True:
        nac = int(rng.choice([16, 32, 64]))
        nwa = (na + nac - 1) // nac
        if nwa > 4:
            continue
        rows = int(rng.choice([4, 8, 16]))
        ny = rows // 4
        if nwa * ny > 4 or (nwa, ny) not in {(1, 1), (2, 1), (3, 1), (4, 1), (1, 2), (2, 2),
                                             (1, 4)}:
            continue
        return nac, rows, int(rng.integers(0, shape[3] + 1))


FP64_SHAPES = _shapes(16, 41)


@pytest.mark.parametrize("shape", FP64_SHAPES, ids=[str(s) for s in FP64_SHAPES])
def test_fp64_random_shape_vs_oracle(shape, cuda):
    import torch
    print(f"case random fp64 {shape}", flush=True)
    d = msw.dims(*shape)
    od = oracle.dims(*shape)
    rng = np.random.default_rng(sum(shape))
    base = oracle.init_state(od)
    vi0 = base * (1.0 + 0.25 * rng.uniform(-1, 1, base.size))
    a, m = oracle.init_matrices(od)
    ref = oracle.sweep(od, a, m, vi0)
    ta, tm = torch.from_numpy(a).to(cuda), torch.from_numpy(m).to(cuda)
    vi = torch.from_numpy(vi0).to(cuda)
    nac, rows, epw = _options(shape, rng)
    runs = [{}, dict(angles_per_wave=nac, tile_rows=rows, groups_per_workgroup=epw)]
    # z blocks: the largest proper divisor of ncell_z, if any
    zb = max([k for k in range(1, shape[2]) if shape[2] % k == 0], default=0)
    if zb:
        runs.append(dict(angles_per_wave=nac, tile_rows=rows, z_block=zb))
    for opts in runs:
        vo = torch.full_like(vi, float("nan"))
        plan = msw.Plan(d, msw.PREC_FP64, engine=msw.ENGINE_MFMA, **opts)
        plan.sweep(vi, vo, ta, tm)
        torch.cuda.synchronize()
        plan.close()
        assert_close(vo.cpu().numpy(), ref, what=f"fp64 {opts}")


MIXED_SHAPES = _shapes(8, 43)


@pytest.mark.parametrize("shape", MIXED_SHAPES, ids=[str(s) for s in MIXED_SHAPES])
def test_mixed_random_shape_manufactured_exact(shape, cuda):
    import torch
    print(f"case random mixed {shape}", flush=True)
    d = msw.dims(*shape)
    a, m = msw.init_matrices(d.nm, d.na)
    ta, tm = torch.from_numpy(a).to(cuda), torch.from_numpy(m).to(cuda)
    vi = torch.empty(msw.size_state(d), dtype=torch.float64, device=cuda)
    msw.init_state(d, vi)
    expect = float(closed_form_normsq(*shape[:5], shape[6]))
    rng = np.random.default_rng(sum(shape))
    nac, rows, epw = _options(shape, rng)
    for prec in (msw.PREC_MIXED_FP32, msw.PREC_MIXED_BF16):
        for opts in ({}, dict(angles_per_wave=nac, tile_rows=rows, groups_per_workgroup=epw)):
            vo = torch.full_like(vi, float("nan"))
            plan = msw.Plan(d, prec, engine=msw.ENGINE_MFMA, **opts)
            plan.sweep(vi, vo, ta, tm)
            torch.cuda.synchronize()
            plan.close()
            assert torch.equal(vo, vi), (prec, opts)
            n, nd = msw.state_norms(d, vi, vo)
            assert n == expect and nd == 0.0
