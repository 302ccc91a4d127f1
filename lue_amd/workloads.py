"""The benchmark workloads beyond the headline step, built from the seeded generators.

  - batch_threshold: accu_threshold over a batch of catchments with sides log-uniform on [lo, hi] and aspect
    U[0.5, 2], each with its own seeded DEM, material U[0, 1) and threshold U[0, 0.5); tile shape autotuned.
  - large_domain: kinematic-wave timesteps on one large raster with the state resident on the device; the
    upstream cell count (int64, exact) sets the initial discharge, lateral inflow is drawn per timestep.
  - serving: kinematic-wave timestep requests (Poisson arrivals) for resident catchments of sides
    log-uniform on [lo, hi] (lue_amd.serving).
"""
from __future__ import annotations

import time

import numpy as np
import torch

from . import autotune
from . import framework as fw
from . import serving
from . import synthetic as syn


def _sync():
    torch.cuda.synchronize()


def d8_flow_direction(seed=20261208, side=16384, repeats=5, device=None):
    """d8_flow_direction on the synthetic DEM (side x side), FP32 and FP64: time per call (HIP events) and
    the HBM rate of its algorithmic bytes (elevation read + ldd written: 5 B/cell FP32, 9 B/cell FP64)."""
    dev = torch.device(device or "cuda")
    result = {"workload": f"d8_flow_direction, {side}x{side} synthetic DEM", "cells": side * side}
    for dtype, name in ((torch.float32, "f32"), (torch.float64, "f64")):
        elevation = fw.PartitionedArray(syn.dem_device(seed, side, side, dtype, dev))
        out = fw.d8_flow_direction(elevation)
        _sync()
        start = torch.cuda.Event(enable_timing=True)
        stop = torch.cuda.Event(enable_timing=True)
        start.record()
        for _ in range(repeats):
            out = fw.d8_flow_direction(elevation)
        stop.record()
        _sync()
        ms = start.elapsed_time(stop) / repeats
        nbytes = side * side * (elevation.data.element_size() + 1)
        result[name] = {"ms": round(ms, 4), "mcells_per_s": round(side * side / ms / 1e3, 1),
                        "algorithmic_bytes": nbytes, "gb_per_s": round(nbytes / ms / 1e6, 1),
                        "hbm_fraction": round(nbytes / ms / 1e6 / 8000.0, 3)}
        del elevation, out
    return result


def batch_threshold(seed=20261205, count=512, side_lo=64, side_hi=2048, repeats=3, device=None):
    dev = torch.device(device or "cuda")
    shapes = syn.batch_shapes(seed, count, side_lo, side_hi)
    ldds, mats, thrs = [], [], []
    for i, (r, c) in enumerate(shapes):
        s = seed + 1 + i
        ldds.append(fw.PartitionedArray(syn.ldd_device(s, r, c, dev)))
        mats.append(fw.PartitionedArray(syn.uniform_device(s, syn.STREAM_MATERIAL, 0.0, 1.0, (r, c), torch.float32, dev)))
        thrs.append(fw.PartitionedArray(syn.uniform_device(s, syn.STREAM_THRESHOLD, 0.0, 0.5, (r, c), torch.float32, dev)))
    cells = sum(r * c for r, c in shapes)
    _sync()
    t0 = time.perf_counter()
    net, table, schedule_table = autotune.tune(ldds, lambda n: n.accu_threshold(mats, thrs), repeats=repeats)
    _sync()
    tune_s = time.perf_counter() - t0
    outs = net.accu_threshold(mats, thrs)
    _sync()
    t0 = time.perf_counter()
    for _ in range(repeats):
        net.accu_threshold(mats, thrs, out=outs[0], remainder=outs[1])
    _sync()
    eager_ms = (time.perf_counter() - t0) / repeats * 1e3
    # the whole batch as one graph launch
    plan = net.plan_accu_threshold(mats, thrs, out=outs[0], remainder=outs[1])
    plan.run()
    _sync()
    t0 = time.perf_counter()
    for _ in range(repeats):
        plan.run()
    _sync()
    ms = (time.perf_counter() - t0) / repeats * 1e3
    launches = plan.info()["launches"]
    plan.close()
    info = net.info()
    net.close()
    return {"workload": f"accu_threshold (float) over {count} catchments, sides log-uniform {side_lo}..{side_hi}, "
                        "aspect U[0.5,2], material U[0,1), threshold U[0,0.5)",
            "catchments": count, "cells": cells, "ms": round(ms, 3), "mcells_per_s": round(cells / ms / 1e3, 1),
            "graph": f"one graph launch of {launches} kernel launches", "eager_ms": round(eager_ms, 3),
            "tile_shape_ms": {f"{k[0]}x{k[1]}": round(v, 3) for k, v in table.items()},
            "wave_schedule_ms": {f"{k[0][0]}x{k[0][1]} wide>={k[1] if k[1] else 'off'}"
                                 f"{' queue' if k[2] else ''}": round(v, 3) for k, v in schedule_table.items()},
            "tile_shape": info["tile_shape"], "wide_threshold": info["wide_threshold"],
            "tile_queue": info["tile_queue"],
            "tune_s": round(tune_s, 2), "levels": info["levels"],
            "pending_cells": info["pending_cells"]}


def upstream_cell_count(ldd, network=None):
    """int64 number of cells upstream of every cell, itself included: accu of 1 in FP64 (exact below 2^53)."""
    ones = fw.PartitionedArray(torch.ones(ldd.shape, dtype=torch.float64, device=ldd.data.device), ldd.partition_shape)
    out = fw.accu(ldd, ones) if network is None else network.accu(ones)
    return torch.nan_to_num(out.data, nan=-1.0).to(torch.int64)


def large_domain(seed=20261206, side=32768, steps=24, partition=1024, dtype=torch.float32, device=None):
    dev = torch.device(device or "cuda")
    shape = (side, side)
    ps = (partition, partition)
    ldd = fw.PartitionedArray(syn.ldd_device(seed, side, side, dev), ps)
    alpha = fw.PartitionedArray(syn.uniform_device(seed, syn.STREAM_ALPHA, 1.5, 2.5, shape, dtype, dev), ps)
    _sync()
    t0 = time.perf_counter()
    net = fw.FlowNetwork(ldd)
    _sync()
    record_s = time.perf_counter() - t0
    counts = upstream_cell_count(ldd, net)
    q = fw.PartitionedArray((counts.to(dtype) * 1e-3), ps)  # initial discharge from the upstream area
    _sync()
    t0 = time.perf_counter()
    for step in range(steps):
        lateral = fw.PartitionedArray(syn.uniform_device(seed + 1 + step, syn.STREAM_LATERAL, 0.0, 1e-4, shape,
                                                         dtype, dev), ps)
        q = net.kinematic_wave(q, lateral, alpha, 0.6, 3600.0, 100.0)
    _sync()
    run_s = time.perf_counter() - t0
    info = net.info()
    total = int(counts[-1].sum())  # every path ends in the bottom row (outlets)
    net.close()
    return {"workload": f"{steps} kinematic_wave timesteps ({'float' if dtype == torch.float32 else 'double'}), "
                        f"{side}x{side} synthetic DEM, {partition}x{partition} partitions, state resident, "
                        "lateral inflow U[0,1e-4) per step, initial discharge 1e-3 x upstream cell count",
            "cells": side * side, "steps": steps, "ms_per_step": round(run_s / steps * 1e3, 3),
            "mcells_per_s": round(side * side * steps / run_s / 1e6, 1), "record_s": round(record_s, 2),
            "outlet_count_sum": total, "levels": info["levels"], "pending_cells": info["pending_cells"],
            "final_discharge_sum": float(q.data.double().sum())}


def serving_run(seed=20261207, count=200, side_lo=256, side_hi=4096, rate=1000.0, duration=10.0,
                budget_gb=200.0, slots=16, device=None):
    dev = torch.device(device or "cuda")
    shapes = syn.batch_shapes(seed, count, side_lo, side_hi)
    server = serving.Server(budget_gb * 1e9, slots=slots, device=dev)
    t0 = time.perf_counter()
    names = []
    for i, (r, c) in enumerate(shapes):
        s = seed + 1 + i
        ldd = fw.PartitionedArray(syn.ldd_device(s, r, c, dev))
        alpha = fw.PartitionedArray(syn.uniform_device(s, syn.STREAM_ALPHA, 1.5, 2.5, (r, c), torch.float32, dev))
        q0 = fw.PartitionedArray(syn.uniform_device(s, syn.STREAM_CURRENT, 0.0, 5.0, (r, c), torch.float32, dev))
        try:
            server.add_catchment(f"c{i}", ldd, alpha, q0)
            names.append(f"c{i}")
        except serving.AdmissionError:
            break
    server.warm_up()
    _sync()
    setup_s = time.perf_counter() - t0
    requests = serving.poisson_requests(seed, rate, duration, names)
    t0 = time.perf_counter()
    server.run(requests)
    _sync()
    wall = time.perf_counter() - t0
    stats = server.statistics(wall)
    cells = sum(r * c for r, c in shapes[:len(names)])
    stats.update({"workload": f"kinematic_wave timestep requests (float), Poisson {rate:g}/s for {duration:g} s, "
                              f"{len(names)} resident catchments of sides log-uniform {side_lo}..{side_hi}, "
                              f"HBM budget {budget_gb:g} GB, {slots} slots",
                  "resident_cells": cells, "setup_s": round(setup_s, 2), "wall_s": round(wall, 2),
                  "offered_requests": len(requests)})
    return stats
