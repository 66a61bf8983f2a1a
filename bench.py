#!/usr/bin/env python3
"""Sweep throughput on one MI355X.

Workload (default): the FP64 plain sweep at one GPU's domain -- 64^3 cells,
ne=64 energy groups, na=64 angles per octant, nm=16 moments, nu=4 unknowns,
8 octants -- with vi/vo (8.6 GB each) resident in HBM and the deterministic
Quantities_init_* input state.  One step = one sweep vo <- S(vi); steps
ping-pong vi/vo as the reference runner does (src/4_driver/runner.c:148-155).
GFLOP/s uses the reference's flop count (src/4_driver/runner.c:162-166).

Prints ONE JSON line.  Also reported in it:
  roofline     dominant kernel (k_sweep_fp64) vs the FP64 peak, from per-launch
               HIP events recorded on the sweep's stream during the timed
               steps (the last timed sweep's launches).
  cpu_baseline the CPU oracle (oracle/, a restatement of the reference
               algorithm) timed on this host over a bounded sample: the same
               64^3 x na64 x nm16 x nu4 shape with a few energy groups, one
               sweep, OpenMP over energy groups.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

METRIC = ("sweep GFLOP/s + p50/p99 latency, mixed vs FP64 plain at ≤1e-5 rel err, "
          "1 GPU")
PEAK_FP64_TFLOPS = 78.6
PEAK_HBM_GBS = 8000.0


def closed_form_normsq(ncx, ncy, ncz, ne, nm, nu):
    s_m = sum((1 + im) ** 2 for im in range(nm))
    s_x = sum(4 ** ((x + 3 * y + 7 * z + 2) & 3)
              for x in range(ncx) for y in range(ncy) for z in range(ncz))
    s_e = sum((1 << (((ie * 1366 + 150889) % 714025) & 3)) ** 2 for ie in range(ne))
    s_u = sum((1 << (((iu * 741 + 66037) % 312500) & 3)) ** 2 for iu in range(nu))
    return s_m * s_x * s_e * s_u


def read_traffic(name):
    """HBM bytes per sweep from the committed rocprofv3 PMC summary, if any."""
    path = os.path.join(REPO, "profiles", name)
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("hbm_bytes_per_sweep:"):
            return float(line.split(":", 1)[1])
    return None


def cpu_baseline(ncx, ncy, ncz, nm, na, nu, threads):
    """Oracle on this host: a few energy groups of the same shape, one sweep."""
    import numpy as np  # noqa: F401
    import oracle
    ne_s = max(1, threads)
    d = oracle.dims(ncx, ncy, ncz, ne_s, nm, na, nu)
    a, m = oracle.init_matrices(d)
    vi = oracle.init_state(d)
    t = time.perf_counter()
    vo = oracle.sweep(d, a, m, vi, threads)
    t = time.perf_counter() - t
    ok = bool((vo == vi).all())
    return {
        "value": round(oracle.flops(d, 1) / t / 1e9, 3),
        "unit": "GFLOP/s",
        "cores": threads,
        "kind": "port",
        "sample": (f"{ncx}x{ncy}x{ncz} cells, ne={ne_s} groups (of the workload's "
                   f"energy groups), na={na}, nm={nm}, nu={nu}, 1 sweep, "
                   f"{t:.1f} s, exact={ok}"),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=4)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--ncell", type=int, default=64)
    ap.add_argument("--ne", type=int, default=64)
    ap.add_argument("--na", type=int, default=64)
    ap.add_argument("--nm", type=int, default=16)
    ap.add_argument("--nu", type=int, default=4)
    ap.add_argument("--engine", choices=["mfma", "valu"], default="mfma")
    ap.add_argument("--precision", choices=["fp64", "fp32", "bf16"], default="fp64")
    ap.add_argument("--angles-per-wave", type=int, default=0)
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="threads for the CPU baseline (0: OMP_NUM_THREADS or 16)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--copy-probe-gib", type=float, default=0.0,
                    help="after the sweeps, one device-to-device copy of this many "
                         "GiB (calibrates rocprofv3 FETCH_SIZE/WRITE_SIZE)")
    ap.add_argument("--launch-times", action="store_true",
                    help="print the last timed sweep's per-launch ms to stderr")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not "
              "supported (the sweep would shard by energy group across GPUs, "
              "see DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    import torch
    import minisweep_amd as msw

    dev = torch.device("cuda:0")
    n = args.ncell
    d = msw.dims(n, n, n, args.ne, args.nm, args.na, args.nu)
    size = msw.size_state(d)
    flops_sweep = msw.flops(d, 1)

    a, m = msw.init_matrices(d.nm, d.na)
    ta, tm = torch.from_numpy(a).to(dev), torch.from_numpy(m).to(dev)
    vi = torch.empty(size, dtype=torch.float64, device=dev)
    vo = torch.empty(size, dtype=torch.float64, device=dev)
    msw.init_state(d, vi)
    engine = msw.ENGINE_MFMA if args.engine == "mfma" else msw.ENGINE_VALU
    prec = {"fp64": msw.PREC_FP64, "fp32": msw.PREC_MIXED_FP32,
            "bf16": msw.PREC_MIXED_BF16}[args.precision]
    plan = msw.Plan(d, prec, engine=engine, angles_per_wave=args.angles_per_wave)
    stream = torch.cuda.current_stream()

    bufs = [vi, vo]
    for i in range(args.warmup):
        plan.sweep(bufs[i % 2], bufs[(i + 1) % 2], ta, tm)
    # restart from the initial state so the timed steps see the same input
    msw.init_state(d, vi)
    torch.cuda.synchronize()

    plan.set_timing(True)
    lat = []
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps + 1)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ev[0].record(stream)
    for i in range(args.steps):
        plan.sweep(bufs[i % 2], bufs[(i + 1) % 2], ta, tm, stream)
        ev[i + 1].record(stream)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for i in range(args.steps):
        lat.append(ev[i].elapsed_time(ev[i + 1]))
    launch_ms = plan.last_timings_ms()
    plan.set_timing(False)

    wall = t1 - t0
    value = flops_sweep * args.steps / wall / 1e9
    kernel_ms = sum(launch_ms)

    # correctness of the timed run: the manufactured problem maps vi to vi
    last_out = bufs[args.steps % 2]
    last_in = bufs[(args.steps + 1) % 2]
    normsq, normsqdiff = msw.state_norms(d, last_in, last_out)
    expect = float(closed_form_normsq(n, n, n, args.ne, args.nm, args.nu))
    exact = normsqdiff == 0.0 and normsq == expect
    normsq_rel_err = abs(normsq - expect) / expect

    lat_sorted = sorted(lat)

    def pct(p):
        k = min(len(lat_sorted) - 1, max(0, int(round(p / 100 * (len(lat_sorted) - 1)))))
        return lat_sorted[k]

    achieved_tf = flops_sweep / (kernel_ms * 1e-3) / 1e12
    traffic = read_traffic("k_sweep_fp64_64cube_traffic.txt")
    roofline = {
        "bound": "mfma",
        "pipe": ("FP64 MFMA 16x16x4 for m2a/a2m, VALU for the solve" if args.engine == "mfma"
                 else "FP64 vector FMA")
                + " (MI355X FP64 vector and matrix peaks are both 78.6 TF)",
        "kernel": "k_sweep_mfma_f64" if args.engine == "mfma" else "k_sweep_fp64",
        "achieved": round(achieved_tf, 3),
        "peak": PEAK_FP64_TFLOPS,
        "unit": "TFLOP/s",
        "frac": round(achieved_tf / PEAK_FP64_TFLOPS, 4),
        "traffic": traffic,
        "launches_per_sweep": len(launch_ms),
        "kernel_ms_per_sweep": round(kernel_ms, 3),
        "compulsory_bytes_per_sweep": 2 * size * 8,
        "compulsory_GBs": round(2 * size * 8 / (kernel_ms * 1e-3) / 1e9, 1),
    }

    cpu = None
    if not args.no_cpu_baseline:
        threads = args.cpu_threads or int(os.environ.get("OMP_NUM_THREADS", "16"))
        threads = max(1, min(threads, 16))
        cpu = cpu_baseline(n, n, n, args.nm, args.na, args.nu, threads)

    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "GFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: deterministic Quantities_init_* state (manufactured problem)",
        "config": {
            "workload": (f"plain FP64 sweep, {n}^3 cells, ne={args.ne}, na={args.na}, "
                         f"nm={args.nm}, nu={args.nu}, 8 octants, state resident in HBM"),
            "ncell": [n, n, n], "ne": args.ne, "na": args.na, "nm": args.nm,
            "nu": args.nu, "precision": args.precision, "engine": args.engine,
            "state_GB_each": round(size * 8 / 1e9, 3),
        },
        "latency_ms": {"p50": round(pct(50), 3), "p99": round(pct(99), 3)},
        "exact": exact,
        "normsq_rel_err": normsq_rel_err,
        "normsq": normsq,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    if args.launch_times:
        print("launch_ms " + " ".join(f"{t:.3f}" for t in launch_ms), file=sys.stderr)
    plan.close()
    if args.copy_probe_gib > 0:
        del vi, vo
        torch.cuda.empty_cache()
        nel = int(args.copy_probe_gib * (1 << 30)) // 8
        src = torch.ones(nel, dtype=torch.float64, device=dev)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        print(f"copy_probe bytes={nel * 8}", file=sys.stderr)


if __name__ == "__main__":
    main()
