#!/usr/bin/env python3
"""Benchmark of the routing hot path: accu + kinematic_wave on a 16384 x 16384 synthetic catchment.

One step = accu(ldd, material) and one kinematic_wave(ldd, Q_cur, q, alpha, beta, dt, dx) timestep over the
whole raster, on one GPU, inputs resident in HBM:
  - ldd from the seeded DEM e(r,c) = (R - r) + 50 |sin(pi c / 256)| + 0.9 u (sqrt(2)-weighted D8,
    bottom row outlets), material U[0,1), Q_cur U[0,5), q U[0,1e-4), alpha 1.5 + U[0,1), beta 0.6,
    dt 3600 s, dx 100 m (the kinematic-wave configuration scaled to 16384^2);
  - value = raster cells per second through both operations, FP32 path (compensated sums, FP32 Newton
    plus one FP64 correction) on flow networks recorded once; the plain FP64 path and the per-call
    entry points are timed beside it.

Prints one JSON line. The roofline is that of the dominant kernel (the tile pass, round 0 of the sweep, of
the kinematic wave): algorithmic bytes per launch / its HIP-event time, and its HBM traffic per launch
from two rocprofv3 --pmc runs of this script (FETCH_SIZE, WRITE_SIZE), calibrated with a copy kernel of
known size. The CPU baseline is the oracle (C++ restatement of LUE's algorithm) on the same inputs on the
host cores.
"""
from __future__ import annotations

import argparse
import csv
import glob
import json
import os
import shutil
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
SEED = 20261203

# algorithmic bytes per cell (compulsory reads + writes)
BYTES_ACCU_F32 = 1 + 4 + 4
BYTES_KW_F32 = 1 + 4 + 4 + 4 + 4
BYTES_ACCU_F64 = 1 + 8 + 8
BYTES_KW_F64 = 1 + 8 + 8 + 8 + 8

KERNELS = {"kinematic_wave_f32": "tile_pass_kernel<3, float", "accu_f32": "tile_pass_kernel<0, float"}
COPY_BYTES = 1 << 30  # calibration copy: 1 GiB read, 1 GiB written


def parse():
    p = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=5)
    p.add_argument("--warmup", type=int, default=2)
    p.add_argument("--size", type=int, default=16384)
    p.add_argument("--fp64-steps", type=int, default=2)
    p.add_argument("--cpu-seconds", type=float, default=10.0,
                   help="minimum CPU-baseline time: the full workload is repeated until it is reached (0: skip)")
    p.add_argument("--cpu-threads", type=int, default=0)
    p.add_argument("--accu-tile", default="auto",
                   help="sweep tile of the accu network (WxH; auto: tile shape, wave schedule and first-round "
                        "launch picked by the autotuner on the FP32 inputs)")
    p.add_argument("--kw-tile", default="auto", help="sweep tile of the kinematic-wave network (WxH or auto)")
    p.add_argument("--schedule", default="chains", choices=["chains", "rounds"])
    p.add_argument("--no-traffic", action="store_true", help="skip the rocprofv3 --pmc traffic runs")
    p.add_argument("--serving-seconds", type=float, default=5.0,
                   help="length of the serving run behind p99_request_ms (200 resident catchments, Poisson "
                        "1000 requests/s; 0: skip; --extras runs it for 10 s)")
    p.add_argument("--extras", action="store_true",
                   help="also run the batch (512 catchments, accu_threshold, autotuned), large-domain (32768^2, "
                        "24 resident timesteps) and serving (200 catchments, Poisson 1000/s) workloads")
    p.add_argument("--probe", action="store_true", help=argparse.SUPPRESS)  # child run under rocprofv3 --pmc
    return p.parse_args()


def log(*a):
    print(*a, file=sys.stderr, flush=True)


# tile shapes of the traffic probes where the bench's own are left to the autotuner: those of the per-call
# entry points (the autotuner picks the same for the kinematic wave at 16384^2)
PROBE_TILES = {"accu": "64x32", "kinematic_wave": "32x32"}


def tile(text, op=None):
    if text == "auto":
        text = PROBE_TILES[op]
    return tuple(int(v) for v in text.split("x"))


def make_inputs(n, dev, torch):
    from lue_amd import framework as fw
    from lue_amd import synthetic as syn

    ps = (1024, 1024) if n >= 1024 else (n, n)
    ldd = fw.PartitionedArray(syn.ldd_device(SEED, n, n, dev), ps)

    def field(stream, lo, hi):
        return fw.PartitionedArray(syn.uniform_device(SEED, stream, lo, hi, (n, n), torch.float32, dev), ps)

    f32 = dict(material=field(syn.STREAM_MATERIAL, 0.0, 1.0), q_cur=field(syn.STREAM_CURRENT, 0.0, 5.0),
               q_lat=field(syn.STREAM_LATERAL, 0.0, 1e-4), alpha=field(syn.STREAM_ALPHA, 1.5, 2.5))
    return ps, ldd, f32


def probe(args):
    """Child run under rocprofv3 --pmc: each operation twice on recorded networks, then the copy kernel."""
    import torch

    from lue_amd import framework as fw

    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    _, ldd, f32 = make_inputs(args.size, dev, torch)
    net_accu = fw.FlowNetwork(ldd, tile_shape=tile(args.accu_tile, "accu"), schedule=args.schedule)
    net_kw = fw.FlowNetwork(ldd, tile_shape=tile(args.kw_tile, "kinematic_wave"), schedule=args.schedule)
    for _ in range(2):
        net_accu.accu(f32["material"])
        net_kw.kinematic_wave(f32["q_cur"], f32["q_lat"], f32["alpha"], 0.6, 3600.0, 100.0)
    torch.cuda.synchronize()
    src = torch.ones(COPY_BYTES // 4, dtype=torch.float32, device=dev)
    dst = torch.empty_like(src)
    for _ in range(2):
        torch.mul(src, 1.0, out=dst)
    torch.cuda.synchronize()
    print("[probe] done", flush=True)


def read_counters(directory, counter):
    """(kernel name, value) per dispatch, in dispatch order, from rocprofv3's counter CSV."""
    files = glob.glob(os.path.join(directory, "**", "*counter_collection.csv"), recursive=True)
    rows = []
    for name in files:
        with open(name, newline="") as f:
            for r in csv.DictReader(f):
                if r.get("Counter_Name") == counter:
                    rows.append((int(r.get("Dispatch_Id", 0)), r["Kernel_Name"], float(r["Counter_Value"])))
    rows.sort()
    return [(k, v) for _, k, v in rows]


def measure_traffic(args):
    """HBM bytes per dispatch of the first-round kernels from two rocprofv3 --pmc runs of --probe."""
    rocprof = shutil.which("rocprofv3")
    if rocprof is None or args.no_traffic:
        return None
    values = {}
    for counter in ("FETCH_SIZE", "WRITE_SIZE"):
        out = tempfile.mkdtemp(prefix="lue_amd_pmc_", dir="/tmp")
        cmd = [rocprof, "--pmc", counter, "--output-format", "csv", "-d", out, "-o", "pmc", "--", sys.executable,
               os.path.abspath(__file__), "--probe", "--size", str(args.size), "--accu-tile", args.accu_tile,
               "--kw-tile", args.kw_tile, "--schedule", args.schedule]
        log(f"[bench] traffic: rocprofv3 --pmc {counter}")
        env = dict(os.environ, TMPDIR="/tmp")
        try:
            run = subprocess.run(cmd, cwd="/tmp", env=env, timeout=300, capture_output=True, text=True)
        except subprocess.TimeoutExpired:
            log(f"[bench] traffic: {counter} run timed out")
            return None
        rows = read_counters(out, counter)
        shutil.rmtree(out, ignore_errors=True)
        if run.returncode != 0 or not rows:
            log(f"[bench] traffic: {counter} run failed (rc {run.returncode}): {run.stderr[-400:]}")
            return None
        copies = [v for k, v in rows if "elementwise" in k]
        if not copies:
            return None
        values[counter] = dict(copy_kib=copies[-1])
        for key, pattern in KERNELS.items():
            found = [v for k, v in rows if pattern in k]
            if found:
                values[counter][key] = found[-1]
    result = {}
    for key in KERNELS:
        if key not in values["FETCH_SIZE"] or key not in values["WRITE_SIZE"]:
            continue
        f_factor = COPY_BYTES / (values["FETCH_SIZE"]["copy_kib"] * 1024.0)
        w_factor = COPY_BYTES / (values["WRITE_SIZE"]["copy_kib"] * 1024.0)
        fetch = values["FETCH_SIZE"][key] * 1024.0 * f_factor
        write = values["WRITE_SIZE"][key] * 1024.0 * w_factor
        result[key] = dict(bytes=fetch + write, fetch_bytes=fetch, write_bytes=write, fetch_factor=f_factor,
                           write_factor=w_factor)
    return result


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "the benchmark runs on one GPU; sharding across GPUs is a design only",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    if args.probe:
        probe(args)
        return

    # traffic runs first, as child processes, before this process touches the GPU
    traffic = measure_traffic(args)

    import torch

    from lue_amd import _lib
    from lue_amd import framework as fw

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    n = args.size
    cells = n * n
    log(f"[bench] generating {n}x{n} inputs on the device")
    ps, ldd, f32 = make_inputs(n, dev, torch)
    f64 = {k: fw.PartitionedArray(v.data.double(), ps) for k, v in f32.items()}
    torch.cuda.synchronize()

    # tile shape, wave schedule and first-round launch of each operation's network: the autotuner times
    # every combination on the FP32 inputs (one network recorded per tile shape)
    settings = {}
    for op, arg in (("accu", args.accu_tile), ("kinematic_wave", args.kw_tile)):
        if arg != "auto":
            settings[op] = dict(tile_shape=tile(arg), wide_threshold=None, tile_queue=None, tuned=False)
            continue
        from lue_amd import autotune

        log(f"[bench] autotuning the {op} network")
        scratch = fw.PartitionedArray(torch.empty((n, n), dtype=torch.float32, device=dev), ps)
        if op == "accu":
            def run(net):
                net.accu(f32["material"], out=scratch)
        else:
            def run(net):
                net.kinematic_wave(f32["q_cur"], f32["q_lat"], f32["alpha"], 0.6, 3600.0, 100.0, out=scratch)
        best, table, _ = autotune.tune(ldd, run, schedule=args.schedule)
        picked = best.info()
        settings[op] = dict(tile_shape=best.tile_shape, wide_threshold=picked["wide_threshold"],
                            tile_queue=bool(picked["tile_queue"]), tuned=True,
                            tile_shape_ms={f"{k[0]}x{k[1]}": round(v, 3) for k, v in table.items()})
        best.close()
        del scratch
    torch.cuda.synchronize()

    # networks: the sweep schedule of the ldd, recorded once
    log("[bench] recording flow networks")
    t0 = time.perf_counter()
    net_accu = fw.FlowNetwork(ldd, tile_shape=settings["accu"]["tile_shape"], schedule=args.schedule)
    net_kw = fw.FlowNetwork(ldd, tile_shape=settings["kinematic_wave"]["tile_shape"], schedule=args.schedule)
    torch.cuda.synchronize()
    t_network = time.perf_counter() - t0
    for net, op in ((net_accu, "accu"), (net_kw, "kinematic_wave")):
        if settings[op]["tuned"]:
            net.set_wide_threshold(settings[op]["wide_threshold"])
            net.set_tile_queue(settings[op]["tile_queue"])
    s_accu = torch.cuda.Stream()
    s_kw = torch.cuda.Stream()

    def step(inp, out=None):
        # accu and kinematic_wave are independent: they run concurrently on two streams
        with torch.cuda.stream(s_accu):
            acc = net_accu.accu(inp["material"], out=None if out is None else out[0])
        with torch.cuda.stream(s_kw):
            q = net_kw.kinematic_wave(inp["q_cur"], inp["q_lat"], inp["alpha"], 0.6, 3600.0, 100.0,
                                      out=None if out is None else out[1])
        return acc, q

    def timed(inp, steps, warmup):
        out = step(inp)
        for _ in range(warmup):
            step(inp, out)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            step(inp, out)
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / steps, out

    # FP32 path ------------------------------------------------------------------------------------
    # the profile records HIP events around each operation's first round on its launch stream (no host
    # synchronisation), so the dominant kernel's duration is measured inside the timed steps themselves
    log(f"[bench] fp32 timed {args.steps} steps")
    out32 = step(f32)
    for _ in range(args.warmup):
        step(f32, out32)
    torch.cuda.synchronize()
    _lib.profile_enable(True)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step(f32, out32)
    torch.cuda.synchronize()
    t32 = (time.perf_counter() - t0) / args.steps
    in_steps = {"accu_f32": _lib.profile_read_op("accu"), "kinematic_wave_f32": _lib.profile_read_op("kinematic_wave")}
    _lib.profile_enable(False)

    # the same step as two graph launches: each operation's launch sequence captured once (plans)
    log("[bench] fp32 graph steps")
    plan_accu = net_accu.plan_accu(f32["material"], out=out32[0])
    plan_kw = net_kw.plan_kinematic_wave(f32["q_cur"], f32["q_lat"], f32["alpha"], 0.6, 3600.0, 100.0, out=out32[1])

    def graph_step():
        with torch.cuda.stream(s_accu):
            plan_accu.run()
        with torch.cuda.stream(s_kw):
            plan_kw.run()

    for _ in range(args.warmup):
        graph_step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        graph_step()
    torch.cuda.synchronize()
    t32_graph = (time.perf_counter() - t0) / args.steps
    launches = {"accu": plan_accu.info()["launches"], "kinematic_wave": plan_kw.info()["launches"]}
    t32_eager = t32
    t32 = min(t32_eager, t32_graph)

    # each operation alone (nothing else on the device): its uncontended first round and rest
    kernel_ms = {}
    for name, op, fn in (("accu_f32", "accu", lambda: net_accu.accu(f32["material"], out=out32[0])),
                         ("kinematic_wave_f32", "kinematic_wave",
                          lambda: net_kw.kinematic_wave(f32["q_cur"], f32["q_lat"], f32["alpha"], 0.6, 3600.0, 100.0,
                                                        out=out32[1]))):
        _lib.profile_enable(True)
        for _ in range(args.steps):
            fn()
            torch.cuda.synchronize()
        alone = _lib.profile_read_op(op)
        _lib.profile_enable(False)
        st = in_steps[name]
        kernel_ms[name] = dict(first_round_ms=st["first_round_ms"] / st["calls"],
                               first_round_min_ms=st["first_round_min_ms"],
                               after_first_round_ms=st["after_first_round_ms"] / st["calls"],
                               alone_first_round_ms=alone["first_round_ms"] / alone["calls"],
                               alone_after_first_round_ms=alone["after_first_round_ms"] / alone["calls"],
                               alone_total_ms=(alone["first_round_ms"] + alone["after_first_round_ms"]) / alone["calls"])

    # drop-in per-call path --------------------------------------------------------------------------
    log("[bench] per-call steps")

    def per_call(inp):
        fw.accu(ldd, inp["material"])
        fw.kinematic_wave(ldd, inp["q_cur"], inp["q_lat"], inp["alpha"], 0.6, 3600.0, 100.0)

    def per_call_time(inp):
        # the median of single steps (each synchronised), so that a one-off stall does not set the figure
        per_call(inp)
        times = []
        for _ in range(max(args.fp64_steps, 3)):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            per_call(inp)
            torch.cuda.synchronize()
            times.append(time.perf_counter() - t0)
        return sorted(times)[len(times) // 2]

    t32_call = per_call_time(f32)

    # cold call: the first accu on an ldd the library has not seen (records its schedule)
    fresh = fw.PartitionedArray(ldd.data.clone(), ps)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    fw.accu(fresh, f32["material"])
    torch.cuda.synchronize()
    cold_ms = (time.perf_counter() - t0) * 1e3
    del fresh

    # FP64 path ------------------------------------------------------------------------------------
    log("[bench] fp64 steps")
    t64, out64 = timed(f64, args.fp64_steps, 1)
    t64_call = per_call_time(f64)

    # accuracy of the FP32 path against the FP64 path on the same inputs
    def rel(a, b):
        a = a.data.double()
        b = b.data
        err = (a - b).abs() / b.abs().clamp_min(1e-30)
        err = torch.nan_to_num(err, nan=0.0)
        return float(err.max()), float(err.mean())

    acc_err = rel(out32[0], out64[0])
    # the FP32 path's scalars are FP32 too (beta = 0.6f): the FP64 reference takes the same value
    kw_ref = net_kw.kinematic_wave(f64["q_cur"], f64["q_lat"], f64["alpha"], float(np.float32(0.6)), 3600.0, 100.0)
    torch.cuda.synchronize()
    kw_err = rel(out32[1], kw_ref)
    del kw_ref
    info_kw = net_kw.info()
    info_accu = net_accu.info()
    levels_kw = net_kw.level_info()

    # accu_info on the same ldd and partition grid (per call: the ldd's network is recorded by the first call)
    log("[bench] accu_info")
    classes = fw.accu_info(ldd)
    torch.cuda.synchronize()
    info_reps = max(3, args.steps)
    t0 = time.perf_counter()
    for _ in range(info_reps):
        classes = fw.accu_info(ldd)
    torch.cuda.synchronize()
    t_info = (time.perf_counter() - t0) / info_reps
    del classes
    accu_info_result = {"ms_per_call": round(t_info * 1e3, 3), "mcells_per_s": round(cells / t_info / 1e6, 1),
                        "partition_shape": list(ps),
                        "note": "per-call accu_info (crossing indicator, accu of it, classes) on the bench's ldd"}

    value = cells / t32 / 1e6
    value64 = cells / t64 / 1e6

    # roofline of the dominant kernel, the first round (tile pass) of one operation: it reads every cell's
    # inputs and writes the outputs of the cells it solves (the others are written by the chains' scatter)
    dominant = max(kernel_ms, key=lambda k: kernel_ms[k]["first_round_ms"])
    info = info_kw if dominant == "kinematic_wave_f32" else info_accu
    solved = cells - info["pending_cells"] - info["no_data_cells"]
    reads, writes = (BYTES_KW_F32 - 4, 4) if dominant == "kinematic_wave_f32" else (BYTES_ACCU_F32 - 4, 4)
    tile_pass_bytes = cells * reads + solved * writes
    dom_ms = kernel_ms[dominant]["first_round_ms"]
    achieved = tile_pass_bytes / (dom_ms * 1e-3) / 1e9
    dom_op = "kinematic_wave" if dominant == "kinematic_wave_f32" else "accu"
    dom_tile = "x".join(str(v) for v in settings[dom_op]["tile_shape"])
    dom_arg = args.kw_tile if dom_op == "kinematic_wave" else args.accu_tile
    # the traffic probes ran on their own networks before the tuning: their figure holds for the same tile
    dom_traffic = traffic.get(dominant) if traffic and tile(dom_arg, dom_op) == tuple(
        settings[dom_op]["tile_shape"]) else None

    # CPU baseline: the oracle on the same inputs, copied to the host --------------------------------
    cpu = None
    if args.cpu_seconds > 0:
        from oracle import oracle

        # the box's CPU share for one GPU is 16 threads
        share = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
        threads = args.cpu_threads or min(16, share)
        log(f"[bench] cpu baseline: oracle on {n}x{n}, {threads} threads, >= {args.cpu_seconds} s")
        host = {k: v.data.cpu().numpy() for k, v in f32.items()}
        ldd_h = ldd.data.cpu().numpy()

        def cpu_rate(rows, nthreads):
            lh = np.ascontiguousarray(ldd_h[n - rows:])
            hs = {k: np.ascontiguousarray(v[n - rows:]) for k, v in host.items()}
            pshape = (min(ps[0], rows), ps[1])
            reps = 0
            t0 = time.perf_counter()
            while True:
                oracle.accu(lh, hs["material"], pshape, np.float32, nthreads)
                oracle.kinematic_wave(lh, hs["q_cur"], hs["q_lat"], hs["alpha"], 0.6, 3600.0, 100.0, pshape,
                                      np.float32, nthreads)
                reps += 1
                tc = time.perf_counter() - t0
                if tc >= args.cpu_seconds:
                    return reps * rows * n / tc / 1e6, reps, tc

        rate, reps, tc = cpu_rate(n, threads)
        band = max(1024, n // 8)
        rate1, reps1, tc1 = cpu_rate(band, 1)
        cpu = {"value": round(rate, 2), "unit": "Mcells/s", "cores": threads, "kind": "port",
               "sample": f"accu + kinematic_wave (float), the same {n}x{n} inputs and {ps[0]}x{ps[1]} partitions, "
                         f"{reps} repetition(s) in {tc:.1f} s; LUE-algorithm CPU restatement (oracle/), "
                         f"partition-parallel std::thread pool; host affinity {share} CPUs",
               "one_thread": {"value": round(rate1, 2), "unit": "Mcells/s", "cores": 1,
                              "sample": f"the bottom {band} rows of the same inputs ({band}x{n} cells), {reps1} "
                                        f"repetition(s) in {tc1:.1f} s on one thread"}}
        del host, ldd_h

    extras = None
    serving_result = None
    if args.extras or args.serving_seconds > 0:
        from lue_amd import workloads

        del f64, out64, out32
        net_accu.close()
        net_kw.close()
        torch.cuda.synchronize()
        torch.cuda.empty_cache()
    if args.serving_seconds > 0 and not args.extras:
        log(f"[bench] serving {args.serving_seconds:g} s")
        serving_result = workloads.serving_run(duration=args.serving_seconds)
        torch.cuda.empty_cache()
    if args.extras:
        extras = {}
        log("[bench] extras: d8_flow_direction 16384^2")
        extras["d8_flow_direction"] = workloads.d8_flow_direction()
        torch.cuda.empty_cache()
        log("[bench] extras: batch of 512 catchments")
        extras["batch_accu_threshold"] = workloads.batch_threshold()
        torch.cuda.empty_cache()
        log("[bench] extras: 32768^2, 24 timesteps")
        extras["large_domain"] = workloads.large_domain()
        torch.cuda.empty_cache()
        log("[bench] extras: serving")
        extras["serving"] = serving_result = workloads.serving_run()
        torch.cuda.empty_cache()

    result = {
        "metric": "Mcells/s accu+kinwave @16384^2 (FP32 path, <=1e-5 rel vs FP64)",
        "value": round(value, 1),
        "unit": "Mcells/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(t32 * 1e3, 3),
        "step_paths_ms": {"graph": round(t32_graph * 1e3, 3), "eager": round(t32_eager * 1e3, 3),
                          "note": "graph: each operation's launch sequence replayed as one HIP graph launch (plan); "
                                  "eager: the same launches issued from the host; value = the faster"},
        "launches_per_step": launches,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"accu + kinematic_wave timestep, {n}x{n} synthetic DEM catchment, "
                               f"{ps[0]}x{ps[1]} partitions, beta 0.6, dt 3600 s, dx 100 m",
                   "cells": cells, "parallelism": "single GPU",
                   "networks": {op: {"tile_shape": "x".join(str(v) for v in st["tile_shape"]),
                                     "wide_threshold": st["wide_threshold"], "tile_queue": st["tile_queue"],
                                     "autotuned": st["tuned"], **({"tile_shape_ms": st["tile_shape_ms"]}
                                                                  if st["tuned"] else {})}
                                for op, st in settings.items()}},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": round(dom_traffic["bytes"]) if dom_traffic else None,
                     "kernel": f"{KERNELS[dominant]}, {dom_tile.replace('x', ', ')}>",
                     "algorithmic_bytes": tile_pass_bytes,
                     "algorithmic_note": f"{reads} B read per cell, {writes} B written per cell the tile pass "
                                         f"solves ({solved} of {cells}); the chains' scatter writes the rest",
                     "kernel_ms": round(dom_ms, 3), "kernel_ms_min": round(kernel_ms[dominant]["first_round_min_ms"], 3),
                     "kernel_ms_note": "HIP events around the kernel on its launch stream over the timed steps "
                                       "(average; the other operation runs concurrently on a second stream)"},
        "cpu_baseline": cpu,
        "fp64_dropin": {"value": round(value64, 1), "unit": "Mcells/s", "ms_per_step": round(t64 * 1e3, 3),
                        "fp32_speedup": round(value / value64, 3),
                        "per_call_ms_per_step": round(t64_call * 1e3, 3)},
        "fp32_per_call": {"value": round(cells / t32_call / 1e6, 1), "ms_per_step": round(t32_call * 1e3, 3),
                          "note": "per-call operations of the Python front end (each enqueues its kernels on the flow network the ldd array keeps)"},
        "network": {"record_ms": round(t_network * 1e3, 1), "kinematic_wave": info_kw, "accu": info_accu,
                    "chain_levels": levels_kw},
        "accu_info": accu_info_result,
        "cold_call_ms": round(cold_ms, 1),
        "accuracy_vs_fp64": {"accu_max_rel": acc_err[0], "accu_mean_rel": acc_err[1],
                             "kinematic_wave_max_rel": kw_err[0], "kinematic_wave_mean_rel": kw_err[1]},
        "kernels": kernel_ms,
        "traffic": traffic,
        "step_hbm_gbs": round(cells * (BYTES_ACCU_F32 + BYTES_KW_F32) / t32 / 1e9, 1),
    }
    if serving_result is not None:
        result["p99_request_ms"] = round(serving_result["p99_ms"], 3)
        result["serving"] = {k: serving_result[k] for k in ("workload", "requests", "p50_ms", "p99_ms", "mean_ms",
                                                            "mcells_per_s", "resident_cells", "wall_s")}
    if extras is not None:
        result["extras"] = extras
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
