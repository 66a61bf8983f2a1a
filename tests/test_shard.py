"""Energy-group sharding on the CPU (minisweep_amd/shard.py): the group split
and, over two gloo ranks, that the ranks' group ranges of the oracle's sweep
(oracle_sweep_range, the reference's per-thread group range
src/3_sweeper/sweeper_kba_c_kernels.h:600-627) are the whole sweep's slabs bit
for bit and that the all-reduced norm partials are the whole problem's norms
(src/2_sweeper_base/array_operations.c:69-110)."""
import os
import socket

import numpy as np
import pytest

import oracle
import minisweep_amd as msw
from minisweep_amd import shard


def test_group_range_partitions():
    for ne in range(1, 21):
        for world in range(1, min(ne, 8) + 1):
            ranges = [shard.group_range(ne, world, r) for r in range(world)]
            assert ranges[0][0] == 0 and ranges[-1][1] == ne
            assert all(a[1] == b[0] for a, b in zip(ranges, ranges[1:]))
            sizes = [e - b for b, e in ranges]
            assert min(sizes) >= 1 and max(sizes) - min(sizes) <= 1
    with pytest.raises(ValueError):
        shard.group_range(3, 4, 0)
    with pytest.raises(ValueError):
        shard.group_range(8, 2, 2)


def test_take_put_groups_round_trip():
    d = msw.dims(3, 2, 4, 5, 4, 16, 2)
    full = np.arange(msw.size_state(d), dtype=np.float64)
    back = np.zeros_like(full)
    for r in range(3):
        b, e = shard.group_range(d.ne, 3, r)
        part = shard.take_groups(full, d, b, e)
        assert part.shape == (d.ncell_z, e - b, d.ncell_y * d.ncell_x * d.nu * d.nm)
        shard.put_groups(back, part, d, b, e)
    assert np.array_equal(back, full)


DIMS = (5, 4, 3, 5, 4, 16, 2)  # ncell x, y, z, ne, nm, na, nu


def _rank(rank, world, port):
    import torch.distributed as dist
    os.environ["MASTER_ADDR"] = "127.0.0.1"
    os.environ["MASTER_PORT"] = str(port)
    dist.init_process_group("gloo", rank=rank, world_size=world)
    try:
        od = oracle.dims(*DIMS)
        d = msw.dims(*DIMS)
        a, m = oracle.init_matrices(od)
        vi = oracle.init_state(od) * (1 + 0.01 * np.random.default_rng(7).standard_normal(
            msw.size_state(d)))
        whole = oracle.sweep(od, a, m, vi)
        b, e = shard.group_range(d.ne, world, rank)
        vo = np.zeros_like(vi)
        oracle.sweep_range(od, a, m, vi, vo, b, e)
        mine = shard.take_groups(vo, d, b, e)
        assert np.array_equal(mine, shard.take_groups(whole, d, b, e)), "share != whole's slab"
        vi_mine = shard.take_groups(vi, d, b, e)
        n, nd = shard.allreduce_norms(float(np.sum(mine * mine)),
                                      float(np.sum((mine - vi_mine) ** 2)))
        n_ref, nd_ref = oracle.norms(od, vi, whole)
        assert abs(n - n_ref) <= 1e-12 * n_ref, (n, n_ref)
        assert abs(nd - nd_ref) <= 1e-12 * nd_ref, (nd, nd_ref)
    finally:
        dist.destroy_process_group()


@pytest.mark.parametrize("world", [2, 3])
def test_ranks_reassemble_the_sweep(world):
    import torch.multiprocessing as mp
    oracle.lib()  # build before the ranks start
    with socket.socket() as s:
        s.bind(("127.0.0.1", 0))
        port = s.getsockname()[1]
    mp.spawn(_rank, args=(world, port), nprocs=world, join=True)
