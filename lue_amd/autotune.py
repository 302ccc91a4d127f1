"""Tuning of flow networks: tile shape, wave schedule and the launch of the first round.

The schedule of a network depends on the ldd and on the tile shape: small tiles leave more cells to the
chain phase, large ones hold fewer tiles per CU. The wave schedule of the chain levels is a threshold on a
level's longest work item (which levels run one work item per workgroup of 16 wavefronts instead of one per
wavefront) and a limit on an item's length (items up to it run one per lane, cell after cell), and the first
round runs one workgroup per tile or as a persistent grid fed by a tile work queue. Which is faster depends
on the catchments (their sizes, the length of their flow paths), so a batch is timed on each candidate with
the operation it will run: one network is recorded per tile shape and replayed with every threshold and
launch, then with every lane limit at the best of those (none needs a new recording).
"""
from __future__ import annotations

import time

import torch

from . import _lib
from .framework import TILE_SHAPES, FlowNetwork

CANDIDATES = ((32, 32), (64, 32), (32, 64), (64, 64))
WIDE_THRESHOLDS = (1024, 2048, 4096, 0)
TILE_QUEUE = (False, True)
LANE_LIMITS = (0, 64, 128)


def _time(run, net, repeats):
    run(net)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(repeats):
        run(net)
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / repeats * 1e3


def tune(flow_directions, run, candidates=CANDIDATES, schedule="chains", repeats=3, wide_thresholds=WIDE_THRESHOLDS,
         tile_queue=TILE_QUEUE, lane_limits=LANE_LIMITS):
    """Record a network per candidate tile shape, time `run(network)` (after one untimed call) for each wide
    threshold and first-round launch, then for each lane limit at the best of those, and return (best network,
    set to its best threshold, launch and lane limit; table of {tile_shape: milliseconds} with each shape's
    best time; table of {(tile_shape, threshold, tile_queue, lane_limit): milliseconds}). Shapes whose
    operation does not fit the tile in LDS are skipped; the other networks are released."""
    best = None
    table = {}
    schedule_table = {}
    for shape in candidates:
        if tuple(shape) not in TILE_SHAPES:
            raise ValueError(f"unknown tile shape {shape}")
        net = FlowNetwork(flow_directions, tile_shape=shape, schedule=schedule)
        try:
            run(net)
        except _lib.RoutingError as error:
            if "does not fit" in str(error):
                net.close()
                continue
            raise
        shape_best = None
        lane = net.info()["lane_limit"]
        for queue in tile_queue or (False,):
            net.set_tile_queue(queue)
            for threshold in wide_thresholds or (net.info()["wide_threshold"],):
                net.set_wide_threshold(threshold)
                ms = _time(run, net, repeats)
                schedule_table[(tuple(shape), threshold, queue, lane)] = ms
                if shape_best is None or ms < shape_best[3]:
                    shape_best = (threshold, queue, lane, ms)
        net.set_wide_threshold(shape_best[0])
        net.set_tile_queue(shape_best[1])
        for limit in lane_limits or ():
            if limit == lane:
                continue
            net.set_lane_limit(limit)
            ms = _time(run, net, repeats)
            schedule_table[(tuple(shape), shape_best[0], shape_best[1], limit)] = ms
            if ms < shape_best[3]:
                shape_best = (shape_best[0], shape_best[1], limit, ms)
        net.set_lane_limit(shape_best[2])
        table[tuple(shape)] = shape_best[3]
        if best is None or shape_best[3] < table[best[0]]:
            if best is not None:
                best[1].close()
            best = (tuple(shape), net)
        else:
            net.close()
    if best is None:
        raise _lib.RoutingError(_lib.STATUS_INVALID_ARGUMENT, "no candidate tile shape fits the operation")
    return best[1], table, schedule_table
