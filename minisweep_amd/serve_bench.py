"""Served sweeps on one MI355X: Poisson arrivals of uneven problems into the
continuous-batching server, against a sequential plain-FP64 server.

Requests: problems drawn as the batched set (minisweep_amd.workloads,
seed 1), manufactured input state built on the device at admission, one sweep
each, FP64 norms returned.  For each arrival rate the same request stream goes
to
  * the served form: FP32-MFMA mixed sweeps, up to --max-active problems in
    flight, admitted by HBM footprint (--budget-gib arena);
  * the baseline: plain FP64 sweeps one at a time (one slot), FIFO.
Latency = completion - submission on the server clock; throughput = requests
/ (last completion - first submission).  Saturation throughput (all requests
submitted at once) is measured for both.  Every result is checked against the
closed-form normsq.

The served form takes its kernel configuration from a batch of the batched
workload's shapes (msw_server_tune_from_batch): by default a batch configured
by hand with --schedule-kind, --stagger, --row-groups and --stream-target
(their defaults are what the autotuner settled on for the batched workload in
its last run; the tuner itself is off, msw_batch.h).  With --tune (and
MSW_BATCH_AUTOTUNE=1 in the environment) that batch of --tune-problems
problems (seed 0) is autotuned first.  --plan-defaults serves with the plan
defaults instead.

    python -m minisweep_amd.serve_bench [--rates 50,200,800] [--requests 10000]

Prints one JSON line per measurement and a summary line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np


def tuned_batch(msw, precision, nprob):
    """An autotuned batch of nprob problems of the batched workload (seed 0)."""
    import torch
    from minisweep_amd.workloads import uneven_problems
    dev = torch.device("cuda:0")
    dims = [msw.dims(*s) for s in uneven_problems(nprob, 0)]
    b = msw.Batch(dims, precision=precision)
    mats, bufs = {}, []
    for i, d in enumerate(dims):
        if (d.nm, d.na) not in mats:
            a, m = msw.init_matrices(d.nm, d.na)
            mats[(d.nm, d.na)] = (torch.from_numpy(a).to(dev), torch.from_numpy(m).to(dev))
        vi = torch.empty(msw.size_state(d), dtype=torch.float64, device=dev)
        msw.init_state(d, vi)
        vo = torch.empty_like(vi)
        bufs.append((vi, vo))
        b.set_problem(i, vi, vo, *mats[(d.nm, d.na)])
    b.autotune(reps=2)
    torch.cuda.synchronize()
    return b, bufs


def run_stream(msw, server_kw, shapes, arrivals, expect, label, progress, tune=None):
    """Submit requests at the arrival times (None: all at once); returns stats."""
    srv = msw.Server(**server_kw)
    if tune is not None:
        srv.tune_from(tune)
    # warm the per-shape caches (schedules, matrices) outside the measurement
    seen = set()
    for i, s in enumerate(shapes):
        k = ((s[0] + 15) // 16, (s[1] + 3) // 4, s[4], s[5])
        if k not in seen:
            seen.add(k)
            srv.submit(msw.dims(*s), -1 - i)
    srv.drain(600)
    srv.poll(1 << 20)
    results = []
    t0 = srv.now()
    last_print = time.perf_counter()
    for i, s in enumerate(shapes):
        if arrivals is not None:
            while True:
                dt = t0 + arrivals[i] - srv.now()
                if dt <= 0:
                    break
                time.sleep(min(dt, 0.002))
        srv.submit(msw.dims(*s), i)
        if (i & 255) == 0:
            results += srv.poll(1 << 16)
        if time.perf_counter() - last_print > 20:
            progress(f"{label}: submitted {i + 1}/{len(shapes)}, done {len(results)}")
            last_print = time.perf_counter()
    while True:
        try:
            srv.drain(15)
            break
        except Exception:
            results += srv.poll(1 << 16)
            progress(f"{label}: draining, done {len(results)}/{len(shapes)}")
    results += srv.poll(1 << 16)
    stats = srv.stats()
    srv.close()
    results = [r for r in results if r["tag"] >= 0]
    assert len(results) == len(shapes), (len(results), len(shapes))
    bad = sum(1 for r in results if r["status"] != 0 or r["normsqdiff"] != 0.0
              or r["normsq"] != expect[r["tag"]])
    lat = np.array([r["t_done"] - r["t_submit"] for r in results])
    first = min(r["t_submit"] for r in results)
    last = max(r["t_done"] for r in results)
    return {
        "label": label,
        "requests": len(results),
        "problems_per_s": round(len(results) / (last - first), 2),
        "latency_ms": {"p50": round(float(np.percentile(lat, 50)) * 1e3, 3),
                       "p99": round(float(np.percentile(lat, 99)) * 1e3, 3),
                       "mean": round(float(lat.mean()) * 1e3, 3)},
        "wrong_results": bad,
        "kernel_launches": stats["kernel_launches"],
        "ticks": stats["ticks"],
        "arena_peak_GB": round(stats["arena_peak_bytes"] / 1e9, 3),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--rates", default="50,200,800")
    ap.add_argument("--requests", type=int, default=10000)
    ap.add_argument("--saturation-requests", type=int, default=2000)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--budget-gib", type=float, default=64.0)
    ap.add_argument("--max-active", type=int, default=256)
    ap.add_argument("--precision", choices=["fp32", "bf16", "fp64"], default="fp32")
    ap.add_argument("--skip-baseline", action="store_true")
    ap.add_argument("--schedule-kind", type=int, default=1)
    ap.add_argument("--stagger", type=int, default=0)
    ap.add_argument("--row-groups", type=int, default=1)
    ap.add_argument("--stream-target", type=int, default=128)
    ap.add_argument("--xcd-map", type=int, default=1)
    ap.add_argument("--plan-defaults", action="store_true")
    ap.add_argument("--tune", action="store_true",
                    help="autotune on the batched workload first (needs MSW_BATCH_AUTOTUNE=1)")
    ap.add_argument("--tune-problems", type=int, default=256)
    args = ap.parse_args()

    import minisweep_amd as msw
    from minisweep_amd.workloads import closed_form_normsq, poisson_arrivals, uneven_problems

    def progress(msg):
        print(f"# {msg}", file=sys.stderr, flush=True)

    prec = {"fp64": msw.PREC_FP64, "fp32": msw.PREC_MIXED_FP32,
            "bf16": msw.PREC_MIXED_BF16}[args.precision]
    served = dict(precision=prec, xcd_map=bool(args.xcd_map),
                  hbm_budget_bytes=int(args.budget_gib * (1 << 30)),
                  max_active=args.max_active)
    seq = dict(precision=msw.PREC_FP64, hbm_budget_bytes=int(args.budget_gib * (1 << 30)),
               max_active=1)
    out = []

    def emit(rec):
        print(json.dumps(rec), flush=True)
        out.append(rec)

    tune, bufs, tuned_desc = None, None, None
    if args.tune:
        progress(f"autotuning the served form on {args.tune_problems} problems")
        tune, bufs = tuned_batch(msw, prec, args.tune_problems)
    elif not args.plan_defaults:
        tune = msw.Batch([msw.dims(*s) for s in uneven_problems(args.tune_problems, 0)],
                         precision=prec, schedule_kind=args.schedule_kind,
                         stagger=bool(args.stagger), row_groups=args.row_groups,
                         stream_target=args.stream_target)
    if tune is not None:
        tuned_desc = tune.describe()
        tuned_desc.pop("tuning", None)

    nsat = args.saturation_requests
    shapes = uneven_problems(nsat, args.seed)
    expect = [float(closed_form_normsq(s[0], s[1], s[2], s[3], s[4], s[6])) for s in shapes]
    sat = run_stream(msw, served, shapes, None, expect, f"served {args.precision} saturation",
                     progress, tune)
    emit(sat)
    sat_seq = None
    if not args.skip_baseline:
        sat_seq = run_stream(msw, seq, shapes, None, expect, "sequential fp64 saturation",
                             progress)
        emit(sat_seq)

    for rate in [float(r) for r in args.rates.split(",") if r]:
        shapes = uneven_problems(args.requests, args.seed)
        expect = [float(closed_form_normsq(s[0], s[1], s[2], s[3], s[4], s[6]))
                  for s in shapes]
        arr = poisson_arrivals(args.requests, rate, args.seed)
        rec = run_stream(msw, served, shapes, arr, expect,
                         f"served {args.precision} at {rate:g} req/s", progress, tune)
        rec["rate"] = rate
        emit(rec)
        if not args.skip_baseline:
            rec = run_stream(msw, seq, shapes, arr, expect, f"sequential fp64 at {rate:g} req/s",
                             progress)
            rec["rate"] = rate
            emit(rec)

    summary = {
        "metric": "served sweeps: saturation problems/s vs sequential plain FP64",
        "served_problems_per_s": sat["problems_per_s"],
        "sequential_fp64_problems_per_s": sat_seq["problems_per_s"] if sat_seq else None,
        "speedup": (round(sat["problems_per_s"] / sat_seq["problems_per_s"], 2)
                    if sat_seq else None),
        "all_correct": all(r["wrong_results"] == 0 for r in out),
        "served_config": tuned_desc,
    }
    if tune is not None:
        tune.close()
    print(json.dumps(summary), flush=True)


if __name__ == "__main__":
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    main()
