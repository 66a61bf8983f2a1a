#!/usr/bin/env python3
"""Sweep throughput on one MI355X.

Workload: the plain FP64 sweep at one GPU's domain -- 64^3 cells, ne=64
energy groups, na=64 angles per octant, nm=16 moments, nu=4 unknowns, 8
octants -- with vi/vo (8.6 GB each) resident in HBM and the deterministic
Quantities_init_* input state.  One step = one sweep vo <- S(vi); steps
ping-pong vi/vo as the reference runner does (src/4_driver/runner.c:148-155).
GFLOP/s uses the reference's flop count (src/4_driver/runner.c:162-166).

Prints ONE JSON line.  `value` is the FP64 sweep's GFLOP/s.  Also in it:
  mixed        the FP32 and BF16 mixed-precision forms on the same workload:
               GFLOP/s, speedup over FP64, normsq error on the manufactured
               input, and the per-output error against the FP64 sweep on a
               perturbed input vi*(1+e), e ~ U(-2^-8, 2^-8) (seed 0).
  roofline     the dominant kernel (k_sweep_mfma) against the FP64 peak, from
               per-launch HIP events recorded on the sweep's stream during the
               timed steps; HBM traffic from the committed rocprofv3 summary.
  cpu_baseline the reference's own CPU sweep, its OpenMP-threads build
               (compiled from its sources into oracle/_ref/ by oracle/Makefile),
               on this host's cores (16 at most), over as many energy groups
               of the same shape, one sweep; beside it its serial build on one
               core and one group ("reference_serial"), and the CPU oracle
               (oracle/, the restatement of the reference algorithm) with
               OpenMP on the same cores ("port").  Without oracle/_ref the
               oracle alone, with the reason in "reference_missing".
  mixed_fp64_last  the mixed FP32 form with a final FP64 sweep: of four
               iterations the first three in FP32, the last in FP64.

With --shards G --shard R it times instead the share of the energy groups
that rank R of G would sweep on its own GPU (a plan with energy_offset, the
whole problem's manufactured state cut to those groups) and prints that one
line.
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
TRAFFIC_FILES = {"fp64": "k_sweep_mfma_f64_64cube_na64.md",
                 "fp32": "k_sweep_mfma_mixed_fp32_64cube_na64.md",
                 "bf16": "k_sweep_mfma_mixed_bf16_64cube_na64.md"}


def closed_form_normsq(ncx, ncy, ncz, ne, nm, nu):
    s_m = sum((1 + im) ** 2 for im in range(nm))
    s_x = sum(4 ** ((x + 3 * y + 7 * z + 2) & 3)
              for x in range(ncx) for y in range(ncy) for z in range(ncz))
    s_e = sum((1 << (((ie * 1366 + 150889) % 714025) & 3)) ** 2 for ie in range(ne))
    s_u = sum((1 << (((iu * 741 + 66037) % 312500) & 3)) ** 2 for iu in range(nu))
    return s_m * s_x * s_e * s_u


def read_traffic(name):
    """(bytes per sweep, bytes per launch) from a committed rocprofv3 summary."""
    path = os.path.join(REPO, "profiles", name)
    if not os.path.exists(path):
        return None, None
    per_sweep = per_launch = None
    for line in open(path):
        if line.startswith("hbm_bytes_per_sweep:"):
            per_sweep = float(line.split(":", 1)[1])
        if line.startswith("hbm_bytes_per_launch:"):
            per_launch = float(line.split(":", 1)[1])
    return per_sweep, per_launch


def cpu_baseline(ncx, ncy, ncz, nm, na, nu, threads):
    """Oracle on this host: a few energy groups of the same shape, one sweep."""
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


def cpu_baseline_reference(ncx, ncy, ncz, nm, na, nu, threads=1):
    """The reference's own sweep driver on this host, one sweep: its serial
    KBA build (oracle/_ref/nm<NM>/sweep) on one energy group, or with
    threads > 1 its OpenMP-threads build (oracle/_ref/nm16_omp/sweep,
    --nthread_e threads) on `threads` groups.  Both are compiled from the
    reference's sources by oracle/Makefile (make ref); its own result line
    gives time and GF/s.  Returns (result, None) or (None, why)."""
    import re
    import subprocess
    if threads > 1:
        exe = os.path.join(REPO, "oracle", "_ref", f"nm{nm}_omp", "sweep")
        ne, extra, build = threads, ["--nthread_e", str(threads)], "OpenMP-threads"
    else:
        exe = os.path.join(REPO, "oracle", "_ref", f"nm{nm}", "sweep")
        ne, extra, build = 1, [], "serial KBA"
    if nu != 4:  # the reference fixes NU = 4
        return None, "the reference is built for nu = 4"
    if not os.access(exe, os.X_OK):
        return None, f"{os.path.relpath(exe, REPO)} is not built (oracle/Makefile ref)"
    cmd = [exe, "--ncell_x", str(ncx), "--ncell_y", str(ncy), "--ncell_z", str(ncz),
           "--ne", str(ne), "--na", str(na), "--niterations", "1"] + extra
    env = dict(os.environ, OMP_NUM_THREADS=str(threads))
    try:
        out = subprocess.run(cmd, capture_output=True, text=True, timeout=300, env=env).stdout
    except (OSError, subprocess.TimeoutExpired) as e:
        return None, f"{os.path.relpath(exe, REPO)} failed: {e}"
    m = re.search(r"Normsq result: (\S+)\s+diff: (\S+)\s+(\w+)\s+time: (\S+)\s+GF/s: (\S+)", out)
    if not m:
        return None, f"{os.path.relpath(exe, REPO)} printed no result line"
    return {
        "value": float(m.group(5)),
        "unit": "GFLOP/s",
        "cores": threads,
        "kind": "reference",
        "sample": (f"reference {build} sweep ({os.path.relpath(exe, REPO)}, built from the "
                   f"reference's sources by oracle/Makefile), {ncx}x{ncy}x{ncz} cells, ne={ne} "
                   f"group(s) of the workload's, na={na}, nm={nm}, nu=4, 1 sweep, "
                   f"{float(m.group(4)):.1f} s, {m.group(3)}"),
    }, None


def pct(vals, p):
    s = sorted(vals)
    k = min(len(s) - 1, max(0, int(round(p / 100 * (len(s) - 1)))))
    return s[k]


def time_precision(msw, torch, d, prec_name, args, ta, tm, vi, vo, flops_sweep):
    """Warmup + exactly `steps` timed sweeps of one precision form."""
    prec = {"fp64": msw.PREC_FP64, "fp32": msw.PREC_MIXED_FP32,
            "bf16": msw.PREC_MIXED_BF16}[prec_name]
    engine = msw.ENGINE_MFMA if args.engine == "mfma" else msw.ENGINE_VALU
    plan = msw.Plan(d, prec, engine=engine, angles_per_wave=args.angles_per_wave,
                    groups_per_workgroup=args.groups_per_workgroup, tile_rows=args.tile_rows)
    stream = torch.cuda.current_stream()
    bufs = [vi, vo]
    for i in range(args.warmup):
        plan.sweep(bufs[i % 2], bufs[(i + 1) % 2], ta, tm, stream)
    msw.init_state(d, vi, stream)  # the timed steps start from the same input
    plan.set_timing(True)
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps + 1)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ev[0].record(stream)
    for i in range(args.steps):
        plan.sweep(bufs[i % 2], bufs[(i + 1) % 2], ta, tm, stream)
        ev[i + 1].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    lat = [ev[i].elapsed_time(ev[i + 1]) for i in range(args.steps)]
    launch_ms = plan.last_timings_ms()
    plan.set_timing(False)
    last_out, last_in = bufs[args.steps % 2], bufs[(args.steps + 1) % 2]
    normsq, normsqdiff = msw.state_norms(d, last_in, last_out)
    return plan, {
        "value": flops_sweep * args.steps / wall / 1e9,
        "ms_per_step": wall / args.steps * 1e3,
        "lat": lat,
        "launch_ms": launch_ms,
        "normsq": normsq,
        "normsqdiff": normsqdiff,
    }


def share_bench(args):
    """One rank's share of the energy groups (--shards G --shard R): the
    groups that rank R of G would sweep on its GPU (minisweep_amd/shard.py),
    swept by a plan whose energy_offset is the share's first group, from the
    whole problem's manufactured state cut to those groups.  Prints one JSON
    line; the G ranks would run such a sweep each, in parallel, plus one
    all-reduce of two values."""
    import torch
    import minisweep_amd as msw
    from minisweep_amd import shard

    dev = torch.device("cuda:0")
    n = args.ncell
    d = msw.dims(n, n, n, args.ne, args.nm, args.na, args.nu)
    ds, off = shard.share_dims(d, args.shards, args.shard)
    whole = torch.empty(msw.size_state(d), dtype=torch.float64, device=dev)
    msw.init_state(d, whole)
    vi0 = shard.take_groups(whole, d, off, off + ds.ne).reshape(-1)
    del whole
    vi, vo = vi0.clone(), torch.empty_like(vi0)
    a, m = msw.init_matrices(d.nm, d.na)
    ta, tm = torch.from_numpy(a).to(dev), torch.from_numpy(m).to(dev)
    name = args.precision or "fp64"
    prec = {"fp64": msw.PREC_FP64, "fp32": msw.PREC_MIXED_FP32,
            "bf16": msw.PREC_MIXED_BF16}[name]
    plan = msw.Plan(ds, prec, angles_per_wave=args.angles_per_wave,
                    groups_per_workgroup=args.groups_per_workgroup, tile_rows=args.tile_rows,
                    energy_offset=off)
    stream = torch.cuda.current_stream()
    bufs = [vi, vo]
    for i in range(args.warmup):
        plan.sweep(bufs[i % 2], bufs[(i + 1) % 2], ta, tm, stream)
    vi.copy_(vi0)
    plan.set_timing(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        plan.sweep(bufs[i % 2], bufs[(i + 1) % 2], ta, tm, stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kernel_ms = sum(plan.last_timings_ms())
    plan.set_timing(False)
    exact = bool(torch.equal(bufs[args.steps % 2], vi0))  # S(vi) = vi, manufactured
    plan.close()
    flops = msw.flops(ds, 1)
    print(json.dumps({
        "metric": "one GPU's share of the sweep (energy groups split over ranks), GFLOP/s",
        "value": round(flops * args.steps / wall / 1e9, 3),
        "unit": "GFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 3),
        "higher_is_better": True,
        "dtype": {"fp64": "f64", "fp32": "f32+bf16 (mixed)", "bf16": "bf16 (mixed)"}[name],
        "data": "synthetic: the whole problem's manufactured state, this share's groups",
        "config": {"workload": (f"share {args.shard} of {args.shards}: energy groups "
                                f"[{off}, {off + ds.ne}) of the {n}^3 ne{args.ne} "
                                f"na{args.na} nm{args.nm} nu{args.nu} sweep, {name}")},
        "share": {"ranks": args.shards, "rank": args.shard, "groups": [off, off + ds.ne]},
        "kernel_ms_per_sweep": round(kernel_ms, 3),
        "exact": exact,
    }))


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
    ap.add_argument("--precision", choices=["fp64", "fp32", "bf16"], default=None,
                    help="time only this form (default: FP64 headline plus both "
                         "mixed forms)")
    ap.add_argument("--angles-per-wave", type=int, default=0)
    ap.add_argument("--groups-per-workgroup", type=int, default=0)
    ap.add_argument("--tile-rows", type=int, default=0)
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="threads for the CPU baseline (0: OMP_NUM_THREADS or 16)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--copy-probe-gib", type=float, default=0.0,
                    help="after the sweeps, one device-to-device copy of this many "
                         "GiB (calibrates rocprofv3 FETCH_SIZE/WRITE_SIZE)")
    ap.add_argument("--shards", type=int, default=1,
                    help="time one rank's share of the energy groups, of this many ranks")
    ap.add_argument("--shard", type=int, default=0, help="the rank whose share --shards times")
    ap.add_argument("--launch-times", action="store_true",
                    help="print the last timed sweep's per-launch ms to stderr")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not "
              "supported (the sweep would shard by energy group across GPUs; "
              "only the one-GPU share is built)", file=sys.stderr)
        sys.exit(2)
    if args.shards > 1:
        share_bench(args)
        return

    # the reference's CPU runs go first, before this process touches the GPU:
    # its OpenMP-threads build on the host's cores (the reported baseline) and
    # its serial build on one core
    ref_cpu = ref_serial = None
    cpu_threads = max(1, min(args.cpu_threads or int(os.environ.get("OMP_NUM_THREADS", "16")),
                             16))
    missing = []
    if not args.no_cpu_baseline:
        ref_cpu, why = cpu_baseline_reference(args.ncell, args.ncell, args.ncell, args.nm,
                                              args.na, args.nu, cpu_threads)
        if why:
            missing.append(why)
        ref_serial, why = cpu_baseline_reference(args.ncell, args.ncell, args.ncell, args.nm,
                                                 args.na, args.nu, 1)
        if why:
            missing.append(why)
        for w in missing:
            print(f"bench.py: cpu_baseline: {w}", file=sys.stderr)

    import torch
    import minisweep_amd as msw

    dev = torch.device("cuda:0")
    n = args.ncell
    d = msw.dims(n, n, n, args.ne, args.nm, args.na, args.nu)
    size = msw.size_state(d)
    flops_sweep = msw.flops(d, 1)
    expect = float(closed_form_normsq(n, n, n, args.ne, args.nm, args.nu))
    default_shape = (n, args.ne, args.na, args.nm, args.nu) == (64, 64, 64, 16, 4)

    a, m = msw.init_matrices(d.nm, d.na)
    ta, tm = torch.from_numpy(a).to(dev), torch.from_numpy(m).to(dev)
    vi = torch.empty(size, dtype=torch.float64, device=dev)
    vo = torch.empty(size, dtype=torch.float64, device=dev)
    msw.init_state(d, vi)

    forms = [args.precision] if args.precision else ["fp64", "fp32", "bf16"]
    head = forms[0]
    res, plans = {}, {}
    for f in forms:
        plans[f], res[f] = time_precision(msw, torch, d, f, args, ta, tm, vi, vo, flops_sweep)

    # per-output error of the mixed forms against the FP64 sweep, full size
    errs = {}
    if len(forms) > 1:
        msw.init_state(d, vi)
        g = torch.Generator(device=dev)
        g.manual_seed(0)
        vi.mul_(1.0 + (torch.rand(size, dtype=torch.float64, device=dev, generator=g)
                       * 2.0 - 1.0) * 2.0 ** -8)
        ref = torch.empty_like(vi)
        plans["fp64"].sweep(vi, ref, ta, tm)
        n_ref, _ = msw.state_norms(d, vi, ref)
        ref_max = float(ref.abs().max())
        for f in forms[1:]:
            plans[f].sweep(vi, vo, ta, tm)
            n_f, _ = msw.state_norms(d, vi, vo)
            vo.sub_(ref)
            errs[f] = {"vo_normwise": float(vo.abs().max()) / ref_max,
                       "normsq_rel": abs(n_f - n_ref) / n_ref}
        torch.cuda.synchronize()

    # The mixed form with a final FP64 sweep: of NIT ping-pong iterations (the
    # reference runner's loop, src/4_driver/runner.c:148-155) the first NIT-1
    # run in mixed FP32 and the last in FP64.  Timed once on the manufactured
    # input (exactness checked).  The error is taken over NIT_ERR = 2
    # iterations of the perturbed input (mixed then FP64, against FP64
    # twice): at 64^3 one sweep takes the perturbation from 2^-8 to values
    # near 1e29 and a second to 6e56, past FP32's range, so two mixed
    # iterations of it are not a measurement.
    final = None
    if len(forms) > 1 and "fp32" in forms:
        nit = 4
        seq = [plans["fp32"]] * (nit - 1) + [plans["fp64"]]

        def chain(plist, x0, out_a, out_b):
            out_a.copy_(x0)
            src, dst = out_a, out_b
            for pl in plist:
                pl.sweep(src, dst, ta, tm)
                src, dst = dst, src
            return src, dst  # (result, its input)

        x0 = vi.clone()  # the perturbed input of the error check above
        msw.init_state(d, vi)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        e0.record()
        res_m, in_m = chain(seq, vi, vo, ref)
        e1.record()
        torch.cuda.synchronize()
        t_chain = e0.elapsed_time(e1)
        n_c, nd_c = msw.state_norms(d, in_m, res_m)
        exact_c = nd_c == 0.0 and n_c == expect
        out64 = torch.empty_like(vi)
        nit_err = 2
        r64, _ = chain([plans["fp64"]] * nit_err, x0, out64, ref)
        r64 = r64.clone()
        n64, _ = msw.state_norms(d, x0, r64)
        r64_max = float(r64.abs().max())
        errs_final = {}
        for name, plist in (("fp64_last", [plans["fp32"]] * (nit_err - 1) + [plans["fp64"]]),):
            rm, _ = chain(plist, x0, vo, ref)
            nm_, _ = msw.state_norms(d, x0, rm)
            errs_final[name] = {"normsq_rel": abs(nm_ - n64) / n64,
                                "vo_normwise": float((rm - r64).abs().max()) / r64_max}
        torch.cuda.synchronize()
        final = {
            "iterations": nit,
            "forms": "fp32 x %d, then fp64 x 1" % (nit - 1),
            "ms_per_run": round(t_chain, 3),
            "value": round(flops_sweep * nit / (t_chain * 1e-3) / 1e9, 3),
            "speedup_vs_fp64": round(nit * res["fp64"]["ms_per_step"] / t_chain, 3),
            "exact_manufactured": bool(exact_c),
            "error_iterations": nit_err,
            "normsq_rel_err_vs_fp64_perturbed": errs_final["fp64_last"]["normsq_rel"],
            "vo_normwise_err_vs_fp64_perturbed": errs_final["fp64_last"]["vo_normwise"],
        }
        del x0, out64, r64

    r = res[head]
    launch_ms = r["launch_ms"]
    kernel_ms = sum(launch_ms)
    achieved_tf = flops_sweep / (kernel_ms * 1e-3) / 1e12
    t_sweep, t_launch = (read_traffic(TRAFFIC_FILES[head])
                         if default_shape and head in TRAFFIC_FILES else (None, None))
    roofline = {
        "bound": "mfma",
        "kernel": f"k_sweep_mfma ({head})" if args.engine == "mfma" else "k_sweep_fp64",
        "achieved": round(achieved_tf, 3),
        "peak": PEAK_FP64_TFLOPS,
        "unit": "TFLOP/s",
        "frac": round(achieved_tf / PEAK_FP64_TFLOPS, 4),
        "traffic": t_launch,
        "traffic_per_sweep": t_sweep,
        "traffic_GBs": (round(t_sweep / (kernel_ms * 1e-3) / 1e9, 1) if t_sweep else None),
        "launches_per_sweep": len(launch_ms),
        "kernel_ms_per_sweep": round(kernel_ms, 3),
        "kernel_avg_launch_ms": round(kernel_ms / max(1, len(launch_ms)), 4),
        "compulsory_bytes_per_sweep": 2 * size * 8,
        "compulsory_GBs": round(2 * size * 8 / (kernel_ms * 1e-3) / 1e9, 1),
        "note": ("achieved = reference flop count of one sweep / summed launch time "
                 "of the sweep kernel (HIP events on the sweep stream); traffic = "
                 "rocprofv3 FETCH_SIZE x2 + WRITE_SIZE x1 (copy-calibrated) per launch"),
    }

    mixed = {}
    for f in forms[1:]:
        mixed[f] = {
            "value": round(res[f]["value"], 3),
            "ms_per_step": round(res[f]["ms_per_step"], 3),
            "speedup_vs_fp64": round(res[f]["value"] / res[head]["value"], 3),
            "latency_ms": {"p50": round(pct(res[f]["lat"], 50), 3),
                           "p99": round(pct(res[f]["lat"], 99), 3)},
            "normsq_rel_err_manufactured": abs(res[f]["normsq"] - expect) / expect,
            "vo_normwise_err_vs_fp64_perturbed": errs[f]["vo_normwise"],
            "normsq_rel_err_vs_fp64_perturbed": errs[f]["normsq_rel"],
        }

    cpu = None
    if not args.no_cpu_baseline:
        port = cpu_baseline(n, n, n, args.nm, args.na, args.nu, cpu_threads)
        if ref_cpu is not None:
            cpu = ref_cpu
            cpu["port"] = port
            cpu["port_over_reference"] = round(port["value"] / ref_cpu["value"], 3)
        else:
            cpu = port
            cpu["reference_missing"] = "; ".join(missing)
        if ref_serial is not None:
            cpu["reference_serial"] = ref_serial

    out = {
        "metric": METRIC,
        "value": round(r["value"], 3),
        "unit": "GFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(r["ms_per_step"], 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if head == "fp64" else f"{head} contractions, f32 faces, f64 state",
        "data": "synthetic: deterministic Quantities_init_* state (manufactured problem)",
        "config": {
            "workload": (f"{'plain FP64' if head == 'fp64' else 'mixed ' + head} sweep, "
                         f"{n}^3 cells, ne={args.ne}, na={args.na}, nm={args.nm}, "
                         f"nu={args.nu}, 8 octants, state resident in HBM"),
            "ncell": [n, n, n], "ne": args.ne, "na": args.na, "nm": args.nm,
            "nu": args.nu, "precision": head, "engine": args.engine,
            "state_GB_each": round(size * 8 / 1e9, 3),
        },
        "latency_ms": {"p50": round(pct(r["lat"], 50), 3), "p99": round(pct(r["lat"], 99), 3)},
        "exact": r["normsqdiff"] == 0.0 and r["normsq"] == expect,
        "normsq_rel_err": abs(r["normsq"] - expect) / expect,
        "mixed": mixed or None,
        "mixed_fp64_last": final,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    if args.launch_times:
        for f in forms:
            print(f"launch_ms {f} " + " ".join(f"{t:.3f}" for t in res[f]["launch_ms"]),
                  file=sys.stderr)
    for p in plans.values():
        p.close()
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
