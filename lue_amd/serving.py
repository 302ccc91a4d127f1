"""Serving kinematic-wave timesteps for resident catchments.

A Server holds catchments on the device: ldd, recorded flow network, alpha and the discharge state. A
request asks for one timestep of one catchment with fresh lateral inflow. Requests are admitted under an
HBM budget (the bytes of the resident catchments plus the transient bytes of the requests in flight must
stay below it) and a number of slots (HIP streams). Admitted requests run as soon as a slot is free, on
that slot's stream, so requests that arrive while others are on the GPU overlap with them (continuous
batching); timesteps of one catchment stay in order because a catchment has at most one request in flight.
A request's latency is from its arrival to the completion of its timestep (HIP event on its stream).

Each timestep runs the catchment's recorded network on its own buffers, so a result does not depend on
which other requests share the GPU (tests check that it equals the timestep run alone).
"""
from __future__ import annotations

import collections
import dataclasses
import gc
import heapq
import time

import numpy as np
import torch

from . import framework as fw
from . import synthetic as syn


@dataclasses.dataclass
class Catchment:
    name: str
    ldd: fw.PartitionedArray
    alpha: fw.PartitionedArray
    state: fw.PartitionedArray        # current discharge, replaced by every timestep
    network: fw.FlowNetwork
    resident_bytes: int
    busy: bool = False
    steps: int = 0


@dataclasses.dataclass(order=True)
class Request:
    arrival: float
    seq: int
    catchment: str = dataclasses.field(compare=False)
    seed: int = dataclasses.field(compare=False)
    lateral: fw.PartitionedArray | None = dataclasses.field(default=None, compare=False)
    # filled when it runs
    admitted: float | None = dataclasses.field(default=None, compare=False)
    done_event: torch.cuda.Event | None = dataclasses.field(default=None, compare=False)
    start_event: torch.cuda.Event | None = dataclasses.field(default=None, compare=False)
    result: fw.PartitionedArray | None = dataclasses.field(default=None, compare=False)
    latency: float | None = dataclasses.field(default=None, compare=False)
    transient_bytes: int = dataclasses.field(default=0, compare=False)


class AdmissionError(RuntimeError):
    pass


class Server:
    """Resident catchments and the request loop. beta, dt and dx are the scalars of every timestep.
    keep_results: keep every request's result (True), none (False), or those of the named catchments."""

    def __init__(self, hbm_budget_bytes, slots=16, beta=0.6, dt=3600.0, dx=100.0, tile_shape=(32, 32),
                 device=None, keep_results=False):
        self.device = torch.device(device or "cuda")
        self.keep_results = keep_results
        self.budget = int(hbm_budget_bytes)
        self.beta, self.dt, self.dx = beta, dt, dx
        self.tile_shape = tile_shape
        self.streams = [torch.cuda.Stream(device=self.device) for _ in range(slots)]
        self.free_slots = list(range(slots))
        self.catchments: dict[str, Catchment] = {}
        self.resident_bytes = 0
        self.transient_bytes = 0
        self.completed: list[Request] = []
        self.cells_done = 0
        self._warm = False

    # -- resident catchments --------------------------------------------------------------------------
    def add_catchment(self, name, ldd, alpha, initial_state):
        """Make a catchment resident: record its network. Raises AdmissionError beyond the HBM budget."""
        rows, cols = ldd.shape
        estimate = rows * cols * (1 + 4 + 4 + 24)  # ldd, alpha, state, recorded schedule (upper estimate)
        if self.resident_bytes + estimate > self.budget:
            raise AdmissionError(f"catchment {name}: {estimate} B exceed the HBM budget")
        network = fw.FlowNetwork(ldd, tile_shape=self.tile_shape)
        self.catchments[name] = Catchment(name, ldd, alpha, initial_state, network, estimate)
        self.resident_bytes += estimate
        self._warm = False
        return self.catchments[name]

    def remove_catchment(self, name):
        c = self.catchments.pop(name)
        if c.busy:
            raise RuntimeError(f"catchment {name} has a request in flight")
        c.network.close()
        self.resident_bytes -= c.resident_bytes

    def warm_up(self):
        """Run one discarded timestep of the largest resident catchment on every slot at once (as many slots
        as the HBM budget admits), so that the memory pools behind the slots' streams hold a request's
        transient buffers before the first request (growing a pool inside the request loop stalls every slot
        for milliseconds). States are unchanged."""
        if not self.catchments:
            return
        c = max(self.catchments.values(), key=lambda x: x.ldd.shape[0] * x.ldd.shape[1])
        need = c.ldd.shape[0] * c.ldd.shape[1] * (4 + 4 + 24)  # as a request: lateral inflow, output, scratch
        slots = min(len(self.streams), max(0, self.budget - self.resident_bytes) // need)
        if slots == 0:
            return
        torch.cuda.synchronize(self.device)
        outs = []
        for stream in self.streams[:slots]:
            with torch.cuda.stream(stream):
                lateral = self.lateral_inflow(0, c.ldd.shape, self.device)
                outs.append((lateral, c.network.kinematic_wave(c.state, lateral, c.alpha, self.beta, self.dt,
                                                               self.dx)))
        torch.cuda.synchronize(self.device)
        del outs
        self._warm = True

    # -- one timestep ------------------------------------------------------------------------------------
    @staticmethod
    def lateral_inflow(seed, shape, device):
        return fw.PartitionedArray(syn.uniform_device(seed, syn.STREAM_LATERAL, 0.0, 1e-4, shape, torch.float32,
                                                      device))

    def _launch(self, request, slot):
        c = self.catchments[request.catchment]
        stream = self.streams[slot]
        with torch.cuda.stream(stream):
            request.start_event = torch.cuda.Event(enable_timing=False)
            request.start_event.record(stream)
            if request.lateral is None:
                request.lateral = self.lateral_inflow(request.seed, c.ldd.shape, self.device)
            out = c.network.kinematic_wave(c.state, request.lateral, c.alpha, self.beta, self.dt, self.dx)
            request.done_event = torch.cuda.Event(enable_timing=False)
            request.done_event.record(stream)
        # the state tensor is read on this stream; keep the old one alive until the step has run
        request.previous_state = c.state
        c.state = out
        request.result = out
        c.busy = True
        c.steps += 1

    def _finish(self, request, slot, now):
        c = self.catchments[request.catchment]
        c.busy = False
        request.latency = now - request.arrival
        self.transient_bytes -= request.transient_bytes
        self.free_slots.append(slot)
        self.cells_done += c.ldd.shape[0] * c.ldd.shape[1]
        # the timestep has run: its inputs can go (the result is the catchment's state now)
        request.lateral = None
        request.previous_state = None
        keep = self.keep_results
        if not (keep is True or (keep and request.catchment in keep)):
            request.result = None
        self.completed.append(request)

    def run(self, requests, clock=time.perf_counter, poll_sleep=50e-6):
        """Serve `requests` (arrival times relative to the start). Returns the completed requests."""
        if not self._warm:
            self.warm_up()
        # the request loop allocates many short-lived objects; a full collection of the cyclic garbage
        # collector over the process's long-lived objects stalls it for tens of milliseconds, so those are
        # frozen and collection is off until the loop ends (request objects are freed by reference counting)
        gc_was_enabled = gc.isenabled()
        gc.freeze()
        gc.disable()
        try:
            return self._loop(requests, clock, poll_sleep)
        finally:
            gc.unfreeze()
            if gc_was_enabled:
                gc.enable()

    def _loop(self, requests, clock, poll_sleep):
        pending = list(requests)
        heapq.heapify(pending)
        waiting = collections.deque()  # arrived, not admitted yet (FIFO)
        in_flight = []  # (request, slot)
        t0 = clock()
        while pending or waiting or in_flight:
            now = clock() - t0
            while pending and pending[0].arrival <= now:
                waiting.append(heapq.heappop(pending))
            # completions
            still = []
            for request, slot in in_flight:
                if request.done_event.query():
                    self._finish(request, slot, clock() - t0)
                else:
                    still.append((request, slot))
            in_flight = still
            # admission: FIFO over the waiting requests whose catchment is idle, while slots and memory last
            kept = collections.deque()
            while waiting:
                request = waiting.popleft()
                c = self.catchments[request.catchment]
                need = c.ldd.shape[0] * c.ldd.shape[1] * (4 + 4 + 24)  # lateral inflow, output, scratch
                if c.busy or not self.free_slots or self.resident_bytes + self.transient_bytes + need > self.budget:
                    kept.append(request)
                    continue
                request.transient_bytes = need
                self.transient_bytes += need
                request.admitted = clock() - t0
                slot = self.free_slots.pop()
                self._launch(request, slot)
                in_flight.append((request, slot))
            waiting = kept
            if not in_flight and waiting:
                # every catchment is idle and every slot free: the request alone exceeds the budget
                raise AdmissionError(f"request for {waiting[0].catchment} does not fit the HBM budget")
            if in_flight or (waiting and not pending):
                time.sleep(poll_sleep)
            elif pending:
                time.sleep(max(0.0, min(poll_sleep, pending[0].arrival - (clock() - t0))))
        return self.completed

    # -- statistics --------------------------------------------------------------------------------------
    def statistics(self, wall_seconds):
        lat = np.array([r.latency for r in self.completed]) * 1e3
        return {"requests": len(self.completed), "p50_ms": float(np.percentile(lat, 50)) if lat.size else None,
                "p99_ms": float(np.percentile(lat, 99)) if lat.size else None,
                "mean_ms": float(lat.mean()) if lat.size else None,
                "mcells_per_s": self.cells_done / wall_seconds / 1e6 if wall_seconds > 0 else None}


def poisson_requests(seed, rate_per_s, duration_s, catchment_names):
    """Requests with exponential inter-arrival times (rate per second) over `duration_s`, each for a
    catchment chosen uniformly; request i carries lateral-inflow seed seed * 1000003 + i."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    i = 0
    while True:
        t += rng.exponential(1.0 / rate_per_s)
        if t > duration_s:
            break
        name = catchment_names[rng.integers(len(catchment_names))]
        out.append(Request(arrival=t, seq=i, catchment=name, seed=seed * 1000003 + i))
        i += 1
    return out
