"""GPU tests of z-blocked sweeps (MswPlanOptions.z_block) and of the
out-of-core sweep (msw_plan_sweep_host), through the C-ABI.

A z-blocked sweep runs octants 0-3 then 4-7, each pass over the z blocks in
its direction, carrying the z faces between blocks; it must match the CPU
oracle like the whole-grid sweep.  The host-streamed sweep runs the same
launches on device slabs that the state is copied through, so its vo is the
same bits as the device sweep of the same plan.
"""
import numpy as np
import pytest

import oracle
import minisweep_amd as msw
from parity import assert_close, closed_form_normsq

pytestmark = pytest.mark.gpu

# (ncx, ncy, ncz, ne, nm, na, nu), z_block
CASES = [
    ((9, 7, 8, 2, 16, 32, 2), 2),
    ((9, 7, 8, 2, 16, 32, 2), 4),
    ((17, 5, 6, 1, 4, 16, 1), 3),
    ((20, 12, 4, 2, 16, 64, 4), 1),
    ((6, 9, 9, 3, 8, 40, 1), 3),
]


def _perturbed(s, seed):
    base = oracle.init_state(oracle.dims(*s))
    rng = np.random.default_rng(seed)
    return base * (1.0 + 0.25 * rng.uniform(-1, 1, base.size))


def _mats(torch, dev, d):
    a, m = msw.init_matrices(d.nm, d.na)
    return torch.from_numpy(a).to(dev), torch.from_numpy(m).to(dev)


@pytest.mark.parametrize("shape,zb", CASES, ids=[f"{c[0]}-zb{c[1]}" for c in CASES])
def test_z_blocks_match_oracle(shape, zb, cuda):
    import torch
    print(f"case z blocks {shape} zb {zb}", flush=True)
    d = msw.dims(*shape)
    vi0 = _perturbed(shape, 3)
    od = oracle.dims(*shape)
    ref = oracle.sweep(od, *oracle.init_matrices(od), vi0)
    ta, tm = _mats(torch, cuda, d)
    vi = torch.from_numpy(vi0).to(cuda)
    vo = torch.full_like(vi, float("nan"))
    plan = msw.Plan(d, msw.PREC_FP64, engine=msw.ENGINE_MFMA, z_block=zb)
    plan.sweep(vi, vo, ta, tm)
    torch.cuda.synchronize()
    plan.close()
    assert_close(vo.cpu().numpy(), ref, what=f"z_block {zb}")


@pytest.mark.parametrize("zb", [0, 2, 4])
def test_host_sweep_matches_device_sweep(zb, cuda):
    """Out-of-core: the same bits as the device sweep of the plan, twice in a
    row (the slabs are reused), and close to the whole-grid sweep."""
    import torch
    print(f"case host sweep zb {zb}", flush=True)
    shape = (13, 10, 8, 3, 16, 32, 2)
    d = msw.dims(*shape)
    vi0 = _perturbed(shape, 5)
    ta, tm = _mats(torch, cuda, d)
    plan = msw.Plan(d, msw.PREC_FP64, engine=msw.ENGINE_MFMA, z_block=zb)
    vi = torch.from_numpy(vi0).to(cuda)
    vo = torch.full_like(vi, float("nan"))
    plan.sweep(vi, vo, ta, tm)
    torch.cuda.synchronize()
    dev = vo.cpu().numpy()
    hvi = torch.from_numpy(vi0).pin_memory()
    for rep in range(2):
        hvo = torch.full((vi0.size,), float("nan"), dtype=torch.float64).pin_memory()
        plan.sweep_host(hvi, hvo, ta, tm)
        torch.cuda.synchronize()
        np.testing.assert_array_equal(hvo.numpy(), dev)
    plan.close()
    whole = msw.Plan(d, msw.PREC_FP64, engine=msw.ENGINE_MFMA)
    vo.fill_(float("nan"))
    whole.sweep(vi, vo, ta, tm)
    torch.cuda.synchronize()
    whole.close()
    assert_close(dev, vo.cpu().numpy(), what="blocked vs whole grid")


def test_host_sweep_manufactured_exact(cuda):
    """A larger out-of-core sweep of the manufactured problem: vo == vi and
    normsq in closed form."""
    import torch
    print("case host sweep manufactured", flush=True)
    shape = (32, 20, 24, 4, 16, 64, 4)
    d = msw.dims(*shape)
    ta, tm = _mats(torch, cuda, d)
    vi = torch.empty(msw.size_state(d), dtype=torch.float64, device=cuda)
    msw.init_state(d, vi)
    hvi = vi.cpu().pin_memory()
    hvo = torch.full_like(hvi, float("nan")).pin_memory()
    plan = msw.Plan(d, msw.PREC_FP64, engine=msw.ENGINE_MFMA, z_block=6)
    plan.sweep_host(hvi, hvo, ta, tm)
    torch.cuda.synchronize()
    plan.close()
    assert torch.equal(hvo, hvi)
    vo = hvo.to(cuda)
    n, nd = msw.state_norms(d, vi, vo)
    assert nd == 0.0 and n == float(closed_form_normsq(*shape[:5], shape[6]))


def test_z_block_options_checked(cuda):
    d = msw.dims(9, 7, 8, 2, 16, 32, 2)
    with pytest.raises(RuntimeError):
        msw.Plan(d, msw.PREC_FP64, engine=msw.ENGINE_MFMA, z_block=3)  # 3 does not divide 8
    with pytest.raises(RuntimeError):
        msw.Plan(d, msw.PREC_MIXED_FP32, engine=msw.ENGINE_MFMA, z_block=4)
    with pytest.raises(RuntimeError):
        msw.Plan(d, msw.PREC_FP64, engine=msw.ENGINE_VALU, z_block=4)


@pytest.mark.parametrize("zb", [2, 4])
def test_host_sweep_matches_oracle(zb, cuda):
    """Out-of-core sweep of a perturbed input against the CPU oracle."""
    import torch
    print(f"case host sweep vs oracle zb {zb}", flush=True)
    shape = (13, 10, 8, 3, 16, 32, 2)
    d = msw.dims(*shape)
    od = oracle.dims(*shape)
    vi0 = _perturbed(shape, 17)
    ta, tm = _mats(torch, cuda, d)
    plan = msw.Plan(d, msw.PREC_FP64, engine=msw.ENGINE_MFMA, z_block=zb)
    hvi = torch.from_numpy(vi0).pin_memory()
    hvo = torch.full((vi0.size,), float("nan"), dtype=torch.float64).pin_memory()
    plan.sweep_host(hvi, hvo, ta, tm)
    torch.cuda.synchronize()
    plan.close()
    assert_close(hvo.numpy(), oracle.sweep(od, *oracle.init_matrices(od), vi0),
                 what=f"host sweep zb {zb} vs oracle")


def test_host_sweeps_ping_pong_without_sync(cuda):
    """Two out-of-core sweeps queued back to back on one stream, the second
    reading the first's output (the reference runner's iteration pattern),
    with no synchronization between them: the same bits as two device
    sweeps.  The second call's copies must wait for the first call's last
    copy back."""
    import torch
    print("case host sweeps ping-pong without sync", flush=True)
    shape = (13, 10, 8, 3, 16, 32, 2)
    d = msw.dims(*shape)
    vi0 = _perturbed(shape, 23)
    ta, tm = _mats(torch, cuda, d)
    plan = msw.Plan(d, msw.PREC_FP64, engine=msw.ENGINE_MFMA, z_block=2)
    a = torch.from_numpy(vi0).to(cuda)
    b = torch.empty_like(a)
    c = torch.empty_like(a)
    plan.sweep(a, b, ta, tm)
    plan.sweep(b, c, ta, tm)
    torch.cuda.synchronize()
    dev1, dev2 = b.cpu().numpy(), c.cpu().numpy()
    h0 = torch.from_numpy(vi0).pin_memory()
    h1 = torch.full((vi0.size,), float("nan"), dtype=torch.float64).pin_memory()
    h2 = torch.full((vi0.size,), float("nan"), dtype=torch.float64).pin_memory()
    plan.sweep_host(h0, h1, ta, tm)
    plan.sweep_host(h1, h2, ta, tm)
    torch.cuda.synchronize()
    plan.close()
    np.testing.assert_array_equal(h1.numpy(), dev1)
    np.testing.assert_array_equal(h2.numpy(), dev2)
