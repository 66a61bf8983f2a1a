"""Batched uneven sweeps on one MI355X: problems/s of one batch run against a
sequential loop of plain FP64 sweeps of the same problems.

Problems: 256 drawn from seed 0 (minisweep_amd.workloads.uneven_problems),
manufactured input state (Quantities_init_state), one sweep each.  The batch
runs the FP32-MFMA mixed form by default, as one hipGraph replay, in the
configuration given by --schedule-kind, --stagger, --row-groups and
--stream-target.  Their defaults are the configuration the autotuner settled
on for this workload in its last run (the reference's octant order, no
stagger, one row group of 4 rows per tile, 128-step streams, the default
angles per wave), set by hand: the tuner itself is off (msw_batch.h).
--autotune, with MSW_BATCH_AUTOTUNE=1 in the environment, runs it first.  The
FP64 batch is timed too, in the library's default configuration, to separate
batching from precision.
Every result is checked: normsq equals the closed form and normsqdiff is 0.

    python -m minisweep_amd.batch_bench [--nprob 256] [--seed 0] [--reps 5]

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
    ap.add_argument("--nprob", type=int, default=256)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--reps", type=int, default=5)
    ap.add_argument("--precision", choices=["fp32", "bf16", "fp64"], default="fp32")
    ap.add_argument("--schedule-kind", type=int, default=1)
    ap.add_argument("--stagger", type=int, default=0)
    ap.add_argument("--row-groups", type=int, default=1)
    ap.add_argument("--stream-target", type=int, default=128)
    ap.add_argument("--xcd-map", type=int, default=1)
    ap.add_argument("--class-nac", action="append", default=[], metavar="NM,NA,NAC",
                    help="angles per wave for the problems with these nm, na in the "
                         "configured batch (repeatable)")
    ap.add_argument("--skip-baselines", action="store_true",
                    help="time only the configured batch (no sequential or FP64 batch runs)")
    ap.add_argument("--autotune", action="store_true",
                    help="autotune the batch first (needs MSW_BATCH_AUTOTUNE=1)")
    args = ap.parse_args()

    import torch
    import minisweep_amd as msw
    from minisweep_amd.workloads import closed_form_normsq, uneven_problems

    dev = torch.device("cuda:0")
    prec = {"fp64": msw.PREC_FP64, "fp32": msw.PREC_MIXED_FP32,
            "bf16": msw.PREC_MIXED_BF16}[args.precision]
    shapes = uneven_problems(args.nprob, args.seed)
    dims = [msw.dims(*s) for s in shapes]
    mats = {}
    vis, vos = [], []
    for d in dims:
        key = (d.nm, d.na)
        if key not in mats:
            a, m = msw.init_matrices(d.nm, d.na)
            mats[key] = (torch.from_numpy(a).to(dev), torch.from_numpy(m).to(dev))
        vi = torch.empty(msw.size_state(d), dtype=torch.float64, device=dev)
        msw.init_state(d, vi)
        vis.append(vi)
        vos.append(torch.empty_like(vi))
    flops = sum(msw.flops(d, 1) for d in dims)
    expect = [float(closed_form_normsq(s[0], s[1], s[2], s[3], s[4], s[6])) for s in shapes]

    def check():
        bad = 0
        for d, vi, vo, e in zip(dims, vis, vos, expect):
            n, nd = msw.state_norms(d, vi, vo)
            bad += int(not (n == e and nd == 0.0))
        return bad

    def time_runs(fn, reps):
        fn()
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(reps):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t) / reps

    # ---- sequential plain FP64: one plan per problem, sweeps one after another
    plans = [msw.Plan(d, msw.PREC_FP64) for d in dims]

    def seq():
        for p, d, vi, vo in zip(plans, dims, vis, vos):
            a, m = mats[(d.nm, d.na)]
            p.sweep(vi, vo, a, m)

    t_seq, bad_seq = float("nan"), 0
    if not args.skip_baselines:
        t_seq = time_runs(seq, max(1, args.reps // 2))
        bad_seq = check()
    for p in plans:
        p.close()

    def batch_of(precision, tune, configured=False):
        kw = dict(schedule_kind=args.schedule_kind, stagger=bool(args.stagger),
                  row_groups=args.row_groups, stream_target=args.stream_target,
                  xcd_map=bool(args.xcd_map)) if configured else {}
        b = msw.Batch(dims, precision=precision, **kw)
        if configured:
            for spec in args.class_nac:
                nm, na, nac = (int(x) for x in spec.split(","))
                b.set_class_nac(nm, na, nac)
        for i, d in enumerate(dims):
            a, m = mats[(d.nm, d.na)]
            b.set_problem(i, vis[i], vos[i], a, m)
        t_tune = None
        if tune:
            t0 = time.perf_counter()
            b.autotune(reps=2)
            t_tune = time.perf_counter() - t0
        return b, t_tune

    t_b64, bad_b64 = float("nan"), 0
    if not args.skip_baselines:
        b64, _ = batch_of(msw.PREC_FP64, False)
        t_b64 = time_runs(b64.run, args.reps)
        bad_b64 = check()
        b64.close()

    bm, t_tune = batch_of(prec, args.autotune, configured=True)
    t_bm = time_runs(bm.run, args.reps)
    bad_bm = check()
    desc = bm.describe()
    ws = bm.workspace_bytes
    bm.close()

    out = {
        "metric": "batched uneven sweeps, problems/s, one hipGraph replay vs sequential FP64",
        "value": round(args.nprob / t_bm, 2),
        "unit": "problems/s",
        "problems": args.nprob,
        "seed": args.seed,
        "precision": args.precision,
        "ms_per_batch": round(t_bm * 1e3, 3),
        "gflops": round(flops / t_bm / 1e9, 1),
        "sequential_fp64": {"problems_per_s": round(args.nprob / t_seq, 2),
                            "ms_total": round(t_seq * 1e3, 3),
                            "gflops": round(flops / t_seq / 1e9, 1)},
        "batched_fp64": {"problems_per_s": round(args.nprob / t_b64, 2),
                         "ms_per_batch": round(t_b64 * 1e3, 3)},
        "speedup_vs_sequential_fp64": round(t_seq / t_bm, 2),
        "speedup_batched_fp64_vs_sequential_fp64": round(t_seq / t_b64, 2),
        "all_exact": bad_seq == 0 and bad_b64 == 0 and bad_bm == 0,
        "autotune_s": round(t_tune, 2) if t_tune else None,
        "tuned": desc,
        "workspace_GB": round(ws / 1e9, 3),
        "state_GB": round(sum(2 * msw.size_state(d) * 8 for d in dims) / 1e9, 3),
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    main()
