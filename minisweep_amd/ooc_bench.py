"""Out-of-core sweeps on one MI355X: the state in pinned host memory, streamed
through HBM one z block at a time (msw_plan_sweep_host), against the same
sweep with the state resident in HBM (whole grid, and z-blocked).

Manufactured input (Quantities_init_state); every result is checked: vo == vi
exactly.  PCIe volume per sweep with n z blocks is 5 - 3/n state volumes:
vi in twice, vo back twice, vo in once, less the block at the turn between
the two passes, which stays on the device.

    python -m minisweep_amd.ooc_bench [--ncell 64] [--ne 64] [--z-block 8] [--reps 2]

Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--ncell", type=int, default=64)
    ap.add_argument("--ne", type=int, default=64)
    ap.add_argument("--na", type=int, default=64)
    ap.add_argument("--nm", type=int, default=16)
    ap.add_argument("--z-block", type=int, default=8)
    ap.add_argument("--reps", type=int, default=2)
    args = ap.parse_args()

    import torch
    import minisweep_amd as msw

    dev = torch.device("cuda:0")
    n = args.ncell
    d = msw.dims(n, n, n, args.ne, args.nm, args.na, 4)
    size = msw.size_state(d)
    nblk = args.ncell // args.z_block if args.z_block else 1
    fl = msw.flops(d, 1)
    a, m = msw.init_matrices(d.nm, d.na)
    ta, tm = torch.from_numpy(a).to(dev), torch.from_numpy(m).to(dev)
    vi = torch.empty(size, dtype=torch.float64, device=dev)
    msw.init_state(d, vi)
    vo = torch.empty_like(vi)

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(args.reps):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t) / args.reps

    out = {"metric": "out-of-core sweep: state in host memory, streamed by z block",
           "config": {"ncell": n, "ne": args.ne, "na": args.na, "nm": args.nm, "nu": 4,
                      "z_block": args.z_block, "state_GB_each": round(size * 8 / 1e9, 2)}}
    whole = msw.Plan(d, msw.PREC_FP64)
    t = timed(lambda: whole.sweep(vi, vo, ta, tm))
    out["resident_whole_grid"] = {"ms": round(t * 1e3, 2), "gflops": round(fl / t / 1e9, 1),
                                  "exact": bool(torch.equal(vo, vi))}
    whole.close()
    blk = msw.Plan(d, msw.PREC_FP64, z_block=args.z_block)
    vo.fill_(0)
    t = timed(lambda: blk.sweep(vi, vo, ta, tm))
    out["resident_z_blocked"] = {"ms": round(t * 1e3, 2), "gflops": round(fl / t / 1e9, 1),
                                 "exact": bool(torch.equal(vo, vi))}
    hvi = vi.cpu().pin_memory()
    del vo
    hvo = torch.zeros_like(hvi).pin_memory()
    t = timed(lambda: blk.sweep_host(hvi, hvo, ta, tm))
    out["host_streamed"] = {"ms": round(t * 1e3, 2), "gflops": round(fl / t / 1e9, 1),
                            "pcie_GBs": round((5 - 3 / nblk) * size * 8 / t / 1e9, 1),
                            "exact": bool(torch.equal(hvo, hvi)),
                            "device_GB": round(blk.workspace_bytes / 1e9, 2)}
    blk.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    main()
