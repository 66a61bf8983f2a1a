"""A share of the energy groups swept on the GPU with a plan whose
energy_offset is the share's first group (minisweep_amd/shard.py): the same
bits as that slab of the whole problem's sweep, in FP64 and in the mixed
forms, exact on the manufactured input, and the shares' norms add up to the
whole problem's.  The groups do not couple (the reference's per-thread
group ranges, src/3_sweeper/sweeper_kba_c_kernels.h:600-627)."""
import numpy as np
import pytest

import oracle
import minisweep_amd as msw
from minisweep_amd import shard
from parity import assert_close

pytestmark = pytest.mark.gpu

PRECS = [(msw.PREC_FP64, "fp64"), (msw.PREC_MIXED_FP32, "fp32"), (msw.PREC_MIXED_BF16, "bf16")]


# ne 7: two ranks start at groups 0 and 4, three at 0, 3 and 5.  The
# manufactured group factor's exponent alternates with the group's parity,
# so an odd first group is what shows a share swept with the wrong offset.
@pytest.mark.parametrize("prec,name", PRECS, ids=[p[1] for p in PRECS])
@pytest.mark.parametrize("world", [2, 3])
def test_shares_match_the_whole_sweep(prec, name, world, cuda):
    import torch
    d = msw.dims(20, 12, 9, 7, 16, 32, 2)
    print(f"case shard {name} world {world}", flush=True)
    a, m = msw.init_matrices(d.nm, d.na)
    ta, tm = torch.from_numpy(a).to(cuda), torch.from_numpy(m).to(cuda)
    rng = np.random.default_rng(11)
    man = torch.empty(msw.size_state(d), dtype=torch.float64, device=cuda)
    msw.init_state(d, man)
    torch.cuda.synchronize()
    inputs = {"manufactured": man,
              "perturbed": man * torch.from_numpy(
                  1 + 0.01 * rng.standard_normal(msw.size_state(d))).to(cuda)}
    for kind, vi in inputs.items():
        vo = torch.empty_like(vi)
        plan = msw.Plan(d, prec)
        plan.sweep(vi, vo, ta, tm)
        torch.cuda.synchronize()
        n_whole, nd_whole = msw.state_norms(d, vi, vo)
        plan.close()
        n_sum = nd_sum = 0.0
        for r in range(world):
            dr, b = shard.share_dims(d, world, r)
            e = b + dr.ne
            vi_r = shard.take_groups(vi, d, b, e).reshape(-1)
            vo_r = torch.full_like(vi_r, float("nan"))
            p = msw.Plan(dr, prec, energy_offset=b)
            p.sweep(vi_r, vo_r, ta, tm)
            torch.cuda.synchronize()
            p.close()
            want = shard.take_groups(vo, d, b, e).reshape(-1)
            assert torch.equal(vo_r, want), f"{kind} rank {r}: share differs from the whole"
            if kind == "manufactured":
                assert torch.equal(vo_r, vi_r), f"rank {r}: manufactured share not exact"
            n, nd = msw.state_norms(dr, vi_r, vo_r)
            n_sum += n
            nd_sum += nd
        assert abs(n_sum - n_whole) <= 1e-12 * n_whole
        assert abs(nd_sum - nd_whole) <= 1e-12 * nd_whole


def test_share_matches_oracle_range(cuda):
    """FP64 share at an offset against the oracle's sweep of the same groups
    of the whole problem, on a well-conditioned perturbed input
    (vi * (1 + U(-1/4, 1/4)), as the random-shape parity tests use)."""
    import torch
    dims_t = (12, 9, 6, 6, 16, 32, 2)
    d, od = msw.dims(*dims_t), oracle.dims(*dims_t)
    print("case shard fp64 vs oracle range", flush=True)
    a, m = oracle.init_matrices(od)
    vi = oracle.init_state(od) * (1 + 0.25 * np.random.default_rng(5).uniform(
        -1, 1, msw.size_state(d)))
    b, e = 3, 5
    ref = np.zeros_like(vi)
    oracle.sweep_range(od, a, m, vi, ref, b, e)
    dr = msw.dims(*dims_t[:3], e - b, *dims_t[4:])
    vi_r = torch.from_numpy(np.ascontiguousarray(shard.take_groups(vi, d, b, e)).reshape(-1)).to(cuda)
    vo_r = torch.empty_like(vi_r)
    p = msw.Plan(dr, msw.PREC_FP64, energy_offset=b)
    p.sweep(vi_r, vo_r, torch.from_numpy(a).to(cuda), torch.from_numpy(m).to(cuda))
    torch.cuda.synchronize()
    p.close()
    assert_close(vo_r.cpu().numpy(), shard.take_groups(ref, d, b, e).reshape(-1),
                 what="share vs oracle range")


def test_negative_energy_offset_rejected(cuda):
    with pytest.raises(msw.SweepError, match="energy_offset"):
        msw.Plan(msw.dims(4, 4, 4, 2, 4, 16, 1), msw.PREC_FP64, energy_offset=-1)
