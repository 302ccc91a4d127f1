#!/usr/bin/env python3
"""Benchmark of the routing hot path: accu + kinematic_wave on a 16384 x 16384 synthetic catchment.

One step = accu(ldd, material) followed by one kinematic_wave(ldd, Q_cur, q, alpha, beta, dt, dx)
timestep over the whole raster, on one GPU, inputs resident in HBM:
  - ldd from the seeded DEM e(r,c) = (R - r) + 50 |sin(pi c / 256)| + 0.9 u (sqrt(2)-weighted D8,
    bottom row outlets), material U[0,1), Q_cur U[0,5), q U[0,1e-4), alpha 1.5 + U[0,1), beta 0.6,
    dt 3600 s, dx 100 m (the kinematic-wave configuration scaled to 16384^2);
  - value = raster cells per second through both operations, FP32 path (compensated sums, FP32 Newton
    plus one FP64 correction); the plain FP64 drop-in is timed beside it.

Prints one JSON line. The CPU baseline is the oracle (C++ restatement of LUE's algorithm), timed on a
bounded sample on the host cores.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
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


def parse():
    p = argparse.ArgumentParser(description=__doc__)
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=5)
    p.add_argument("--warmup", type=int, default=2)
    p.add_argument("--size", type=int, default=16384)
    p.add_argument("--fp64-steps", type=int, default=2)
    p.add_argument("--cpu-seconds", type=float, default=10.0,
                   help="minimum CPU-baseline time: the full workload is repeated until it is reached (0: skip)")
    p.add_argument("--cpu-threads", type=int, default=0)
    p.add_argument("--accu-tile", default="32x32", help="sweep tile of the accu network (WxH)")
    p.add_argument("--kw-tile", default="32x32", help="sweep tile of the kinematic-wave network (WxH)")
    p.add_argument("--schedule", default="chains", choices=["chains", "rounds"])
    return p.parse_args()


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU sharding is design-only "
                                   "(see DESIGN.md)", "n_gpus": args.gpus}))
        sys.exit(2)

    import torch

    from lue_amd import _lib
    from lue_amd import framework as fw
    from lue_amd import synthetic as syn

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    n = args.size
    ps = (1024, 1024) if n >= 1024 else (n, n)
    shape = (n, n)
    cells = n * n

    log(f"[bench] generating {n}x{n} inputs on the device")
    ldd = fw.PartitionedArray(syn.ldd_device(SEED, n, n, dev), ps)

    def field(stream, lo, hi, dtype):
        return fw.PartitionedArray(syn.uniform_device(SEED, stream, lo, hi, shape, dtype, dev), ps)

    f32 = dict(material=field(syn.STREAM_MATERIAL, 0.0, 1.0, torch.float32),
               q_cur=field(syn.STREAM_CURRENT, 0.0, 5.0, torch.float32),
               q_lat=field(syn.STREAM_LATERAL, 0.0, 1e-4, torch.float32),
               alpha=field(syn.STREAM_ALPHA, 1.5, 2.5, torch.float32))
    f64 = {k: fw.PartitionedArray(v.data.double(), ps) for k, v in f32.items()}
    torch.cuda.synchronize()

    # networks: the sweep schedule of the ldd, recorded once (tile shape per operation)
    log("[bench] recording flow networks")
    t0 = time.perf_counter()
    accu_tile = tuple(int(v) for v in args.accu_tile.split("x"))
    kw_tile = tuple(int(v) for v in args.kw_tile.split("x"))
    net_accu = fw.FlowNetwork(ldd, tile_shape=accu_tile, schedule=args.schedule)
    net_kw = fw.FlowNetwork(ldd, tile_shape=kw_tile, schedule=args.schedule)
    torch.cuda.synchronize()
    t_network = time.perf_counter() - t0
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

    def sync():
        s_accu.synchronize()
        s_kw.synchronize()

    def timed(inp, steps, warmup):
        out = step(inp)
        for _ in range(warmup):
            step(inp, out)
        sync()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            step(inp, out)
        sync()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / steps, out

    # FP32 path -----------------------------------------------------------------------------------
    log(f"[bench] fp32 timed {args.steps} steps")
    t32, out32 = timed(f32, args.steps, args.warmup)

    # per-kernel timing of each operation with HIP events on its launch stream
    kernel_ms = {}
    for name, fn in (("accu_f32", lambda: net_accu.accu(f32["material"], out=out32[0])),
                     ("kinematic_wave_f32", lambda: net_kw.kinematic_wave(f32["q_cur"], f32["q_lat"], f32["alpha"],
                                                                          0.6, 3600.0, 100.0, out=out32[1]))):
        _lib.profile_enable(True)
        for _ in range(args.steps):
            fn()
        torch.cuda.synchronize()
        first, rest, calls, rounds, _ = _lib.profile_read()
        _lib.profile_enable(False)
        kernel_ms[name] = dict(first_round_ms=first / calls, after_first_round_ms=rest / calls,
                               total_ms=(first + rest) / calls, rounds=rounds / calls)

    # drop-in per-call path (records its schedule in every call)
    log("[bench] per-call fp32 step")
    fw.accu(ldd, f32["material"])
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.fp64_steps):
        fw.accu(ldd, f32["material"])
        fw.kinematic_wave(ldd, f32["q_cur"], f32["q_lat"], f32["alpha"], 0.6, 3600.0, 100.0)
    torch.cuda.synchronize()
    t32_call = (time.perf_counter() - t0) / args.fp64_steps

    # FP64 drop-in --------------------------------------------------------------------------------
    log("[bench] fp64 steps")
    t64, out64 = timed(f64, args.fp64_steps, 0)
    t0 = time.perf_counter()
    for _ in range(args.fp64_steps):
        fw.accu(ldd, f64["material"])
        fw.kinematic_wave(ldd, f64["q_cur"], f64["q_lat"], f64["alpha"], 0.6, 3600.0, 100.0)
    torch.cuda.synchronize()
    t64_call = (time.perf_counter() - t0) / args.fp64_steps
    info_kw = net_kw.info()
    info_accu = net_accu.info()

    # accuracy of the FP32 path against the FP64 drop-in on the same inputs
    def rel(a, b):
        a = a.data.double()
        b = b.data
        err = (a - b).abs() / b.abs().clamp_min(1e-30)
        err = torch.nan_to_num(err, nan=0.0)
        return float(err.max()), float(err.mean())

    acc_err = rel(out32[0], out64[0])
    kw_err = rel(out32[1], out64[1])

    value = cells / t32 / 1e6
    value64 = cells / t64 / 1e6

    # roofline of the dominant kernel: the all-tile round of the kinematic-wave sweep
    kw = kernel_ms["kinematic_wave_f32"]
    ac = kernel_ms["accu_f32"]
    dominant = "kinematic_wave_f32" if kw["first_round_ms"] >= ac["first_round_ms"] else "accu_f32"
    dom_bytes = cells * (BYTES_KW_F32 if dominant == "kinematic_wave_f32" else BYTES_ACCU_F32)
    dom_ms = kernel_ms[dominant]["first_round_ms"]
    achieved = dom_bytes / (dom_ms * 1e-3) / 1e9

    # CPU baseline: the oracle on the same inputs, copied to the host ---------------------------------
    cpu = None
    if args.cpu_seconds > 0:
        from oracle import oracle

        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        log(f"[bench] cpu baseline: oracle on {n}x{n}, {threads} threads, >= {args.cpu_seconds} s")
        host = {k: v.data.cpu().numpy() for k, v in f32.items()}
        ldd_h = ldd.data.cpu().numpy()
        reps = 0
        t0 = time.perf_counter()
        while True:
            oracle.accu(ldd_h, host["material"], ps, np.float32, threads)
            oracle.kinematic_wave(ldd_h, host["q_cur"], host["q_lat"], host["alpha"], 0.6, 3600.0, 100.0, ps,
                                  np.float32, threads)
            reps += 1
            tc = time.perf_counter() - t0
            if tc >= args.cpu_seconds:
                break
        cpu = {"value": round(reps * cells / tc / 1e6, 2), "unit": "Mcells/s", "cores": threads, "kind": "port",
               "sample": f"accu + kinematic_wave (float), the same {n}x{n} inputs and {ps[0]}x{ps[1]} partitions, "
                         f"{reps} repetition(s) in {tc:.1f} s; LUE-algorithm CPU restatement (oracle/), "
                         f"partition-parallel std::thread pool, host has {os.cpu_count()} logical CPUs"}
        del host, ldd_h

    result = {
        "metric": "Mcells/s accu+kinwave @16384^2 (FP32 path, <=1e-5 rel vs FP64)",
        "value": round(value, 1),
        "unit": "Mcells/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(t32 * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"accu + kinematic_wave timestep, {n}x{n} synthetic DEM catchment, "
                               f"{ps[0]}x{ps[1]} partitions, beta 0.6, dt 3600 s, dx 100 m",
                   "cells": cells, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": f"sweep_first_round ({dominant})"},
        "cpu_baseline": cpu,
        "fp64_dropin": {"value": round(value64, 1), "unit": "Mcells/s", "ms_per_step": round(t64 * 1e3, 3),
                        "fp32_speedup": round(value / value64, 3),
                        "per_call_ms_per_step": round(t64_call * 1e3, 3)},
        "fp32_per_call": {"value": round(cells / t32_call / 1e6, 1), "ms_per_step": round(t32_call * 1e3, 3),
                          "note": "drop-in entry points, schedule recorded inside every call"},
        "network": {"record_ms": round(t_network * 1e3, 1), "kinematic_wave": info_kw, "accu": info_accu},
        "accuracy_vs_fp64": {"accu_max_rel": acc_err[0], "accu_mean_rel": acc_err[1],
                             "kinematic_wave_max_rel": kw_err[0], "kinematic_wave_mean_rel": kw_err[1]},
        "kernels": kernel_ms,
        "step_hbm_gbs": round(cells * (BYTES_ACCU_F32 + BYTES_KW_F32) / t32 / 1e9, 1),
    }
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
