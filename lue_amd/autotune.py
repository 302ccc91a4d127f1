"""Tile-shape tuning of flow networks: record one network per candidate tile shape and keep the fastest.

The schedule of a network depends on the ldd and on the tile shape: small tiles leave more cells to the
chain phase, large ones hold fewer tiles per CU. Which is faster depends on the catchments (their sizes,
the length of their flow paths), so a batch is timed on each candidate with the operation it will run.
"""
from __future__ import annotations

import time

import torch

from . import _lib
from .framework import TILE_SHAPES, FlowNetwork

CANDIDATES = ((32, 32), (64, 32), (64, 64))


def tune(flow_directions, run, candidates=CANDIDATES, schedule="chains", repeats=3):
    """Record a network per candidate tile shape, time `run(network)` (after one untimed call) and return
    (best network, table of {tile_shape: milliseconds}). Shapes whose operation does not fit the tile in
    LDS are skipped; the other networks are released."""
    best = None
    table = {}
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
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(repeats):
            run(net)
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) / repeats * 1e3
        table[tuple(shape)] = ms
        if best is None or ms < table[best[0]]:
            if best is not None:
                best[1].close()
            best = (tuple(shape), net)
        else:
            net.close()
    if best is None:
        raise _lib.RoutingError(_lib.STATUS_INVALID_ARGUMENT, "no candidate tile shape fits the operation")
    return best[1], table
