"""Energy-group sharding of one sweep over several processes, one per GPU.

A sweep's energy groups do not couple: faces, solve, m2a and a2m are all
per group, and the reference itself hands each thread a contiguous range of
groups (iemin/iemax, src/3_sweeper/sweeper_kba_c_kernels.h:600-627); the
per-group scale factor of the manufactured problem depends only on the
global group index (src/3_sweeper/quantities_testing_kernels.h:52-70).  So a problem of ne
groups splits into contiguous group ranges, one per rank.  A rank keeps the
[z][its groups] slabs of vi and vo as a state of its own ne_r groups, in the
usual layout, and sweeps it with a plan whose ``energy_offset`` is its first
group; at the same angles per wave its vo is, bit for bit, that slab of the
whole problem's vo (a plan's FP64 default depends on the problem's size, so
a share can default to 16 angles per wave where the whole takes 32; pass
``angles_per_wave`` to pin it).  The only
exchange per sweep is the norms: each rank's FP64 normsq and normsqdiff
partials, summed by one all-reduce of two values (the reference's
get_state_norms over all groups, src/4_driver/runner.c:173-174 and
src/2_sweeper_base/array_operations.c:69-110).
"""
from __future__ import annotations

from typing import Tuple

from . import MswDims, dims


def group_range(ne: int, world: int, rank: int) -> Tuple[int, int]:
    """Groups [begin, end) of rank ``rank`` of ``world``: contiguous, sizes
    differing by at most one, the first ne % world ranks one larger."""
    if world < 1 or not 0 <= rank < world:
        raise ValueError(f"rank {rank} of world {world}")
    if world > ne:
        raise ValueError(f"{world} ranks for {ne} energy groups: every rank needs one")
    base, extra = divmod(ne, world)
    begin = rank * base + min(rank, extra)
    return begin, begin + base + (1 if rank < extra else 0)


def share_dims(d: MswDims, world: int, rank: int) -> Tuple[MswDims, int]:
    """The dims of rank ``rank``'s share of problem ``d`` and its first group
    (the plan's ``energy_offset``)."""
    b, e = group_range(d.ne, world, rank)
    return dims(d.ncell_x, d.ncell_y, d.ncell_z, e - b, d.nm, d.na, d.nu), b


def take_groups(state, d: MswDims, begin: int, end: int):
    """Groups [begin, end) of a whole-problem state (numpy array or torch
    tensor in the [z][group][y][x][unknown][moment] layout), as a contiguous
    state of end - begin groups."""
    v = state.reshape(d.ncell_z, d.ne, -1)[:, begin:end]
    return v.contiguous() if hasattr(v, "contiguous") else v.copy()


def put_groups(state, part, d: MswDims, begin: int, end: int) -> None:
    """Write a share's state ``part`` (groups [begin, end)) into the
    whole-problem state ``state`` in place."""
    state.reshape(d.ncell_z, d.ne, -1)[:, begin:end] = part.reshape(d.ncell_z, end - begin, -1)


def allreduce_norms(normsq: float, normsqdiff: float, group=None) -> Tuple[float, float]:
    """Sum the ranks' norm partials: one all-reduce of two FP64 values over
    torch.distributed's default (or the given) process group.  On gloo the
    values travel in host memory; on nccl (RCCL) in the current GPU's."""
    import torch
    import torch.distributed as dist

    device = "cpu"
    if dist.get_backend(group) == "nccl":
        device = torch.device("cuda", torch.cuda.current_device())
    t = torch.tensor([normsq, normsqdiff], dtype=torch.float64, device=device)
    dist.all_reduce(t, op=dist.ReduceOp.SUM, group=group)
    t = t.cpu()
    return float(t[0]), float(t[1])
