"""The reference's octant order as a tile schedule, and oracle checks of the
batched and served forms on non-affine input.

schedule_kind 1 puts the octants in the KBA order {0,4,2,6,3,7,1,5}
(src/3_sweeper/stepscheduler_kba.c:155) ahead of each other when a launch's
(octant, tile) visits are picked; kind 0 (the default) favours the octant
with the most visits left.  Either way every physical tile gets its eight
visits and every octant sweeps upwind first, so the result is the sweep up to
the order in which the octants' contributions are summed: both are held to
the oracle at the FP64 bar, on whole-grid plans, z-blocked plans, batches
and served requests.
"""
import numpy as np
import pytest

import oracle
import minisweep_amd as msw
from parity import assert_close

pytestmark = pytest.mark.gpu

SHAPES = [
    (9, 7, 3, 2, 16, 64, 4),
    (17, 3, 5, 1, 4, 32, 2),
    (33, 20, 4, 2, 16, 32, 2),
    (5, 40, 6, 1, 8, 48, 1),
]


def _perturbed(s, seed):
    base = oracle.init_state(oracle.dims(*s))
    rng = np.random.default_rng(seed)
    return base * (1.0 + 0.25 * rng.uniform(-1, 1, base.size))


def _ref(s, vi0):
    od = oracle.dims(*s)
    return oracle.sweep(od, *oracle.init_matrices(od), vi0)


def _mats(torch, dev, d):
    a, m = msw.init_matrices(d.nm, d.na)
    return torch.from_numpy(a).to(dev), torch.from_numpy(m).to(dev)


@pytest.mark.parametrize("kind", [0, 1])
@pytest.mark.parametrize("shape", SHAPES, ids=[str(s) for s in SHAPES])
def test_plan_schedule_kinds_match_oracle(shape, kind, cuda):
    import torch
    print(f"case plan {shape} schedule kind {kind}", flush=True)
    d = msw.dims(*shape)
    vi0 = _perturbed(shape, 31)
    ta, tm = _mats(torch, cuda, d)
    vi = torch.from_numpy(vi0).to(cuda)
    vo = torch.full_like(vi, float("nan"))
    plan = msw.Plan(d, schedule_kind=kind)
    plan.sweep(vi, vo, ta, tm)
    torch.cuda.synchronize()
    plan.close()
    assert_close(vo.cpu().numpy(), _ref(shape, vi0), what=f"kind {kind}")


def test_z_blocked_reference_order_matches_oracle(cuda):
    import torch
    shape = (9, 7, 8, 2, 16, 32, 2)
    print(f"case z-blocked {shape} reference order", flush=True)
    d = msw.dims(*shape)
    vi0 = _perturbed(shape, 37)
    ta, tm = _mats(torch, cuda, d)
    vi = torch.from_numpy(vi0).to(cuda)
    vo = torch.full_like(vi, float("nan"))
    plan = msw.Plan(d, z_block=4, schedule_kind=1)
    plan.sweep(vi, vo, ta, tm)
    torch.cuda.synchronize()
    plan.close()
    assert_close(vo.cpu().numpy(), _ref(shape, vi0), what="z-blocked, reference order")


def test_schedule_kind_checked(cuda):
    with pytest.raises(RuntimeError):
        msw.Plan(msw.dims(4, 4, 4, 1, 4, 16, 1), schedule_kind=2)


@pytest.mark.parametrize("kind", [0, 1])
def test_batch_schedule_kinds_match_oracle(kind, cuda):
    import torch
    print(f"case batch schedule kind {kind}", flush=True)
    ins = []
    for i, s in enumerate(SHAPES):
        d = msw.dims(*s)
        vi0 = _perturbed(s, 41 + i)
        ta, tm = _mats(torch, cuda, d)
        vi = torch.from_numpy(vi0).to(cuda)
        ins.append((s, d, vi0, vi, torch.full_like(vi, float("nan")), ta, tm))
    b = msw.Batch([x[1] for x in ins], precision=msw.PREC_FP64, schedule_kind=kind)
    for i, x in enumerate(ins):
        b.set_problem(i, x[3], x[4], x[5], x[6])
    b.run()
    torch.cuda.synchronize()
    for s, d, vi0, vi, vo, ta, tm in ins:
        assert_close(vo.cpu().numpy(), _ref(s, vi0), what=f"batch kind {kind} {s}")
    b.close()


@pytest.mark.parametrize("kind", [0, 1])
def test_served_fp64_perturbed_matches_oracle(kind, cuda):
    """FP64 requests with host-supplied non-affine input, several in flight
    at once, each held to the oracle."""
    print(f"case server fp64 perturbed, schedule kind {kind}", flush=True)
    srv = msw.Server(precision=msw.PREC_FP64, hbm_budget_bytes=1 << 30, max_active=8,
                     keep_outputs=True, schedule_kind=kind)
    inputs = {}
    for t, s in enumerate(SHAPES):
        inputs[t] = _perturbed(s, 53 + t)
        srv.submit(msw.dims(*s), t, inputs[t])
    srv.drain(60)
    res = srv.poll()
    assert sorted(r["tag"] for r in res) == list(range(len(SHAPES)))
    assert all(r["status"] == 0 for r in res)
    for t, s in enumerate(SHAPES):
        got = srv.fetch(t, msw.dims(*s))
        ref = _ref(s, inputs[t])
        assert_close(got, ref, what=f"served {s}")
        r = [x for x in res if x["tag"] == t][0]
        n_ref, nd_ref = oracle.norms(oracle.dims(*s), inputs[t], ref)
        assert abs(r["normsq"] - n_ref) <= 1e-12 * n_ref
        assert abs(r["normsqdiff"] - nd_ref) <= 1e-12 * nd_ref
    srv.close()
