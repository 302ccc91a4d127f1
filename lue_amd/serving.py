"""Serving kinematic-wave timesteps for resident catchments.

A Server holds catchments on the device: ldd, recorded flow network, alpha, the discharge state (two
buffers, the current one and the next), a lateral-inflow buffer and the catchment's timestep captured twice
as HIP graphs (plans: current -> next, and back). Catchments are admitted under an HBM budget for all of
that. A request asks for one timestep of one catchment with fresh lateral inflow: it writes the inflow into
the catchment's buffer and launches the catchment's graph, so it allocates nothing and costs two launches.
Requests run as soon as a slot (HIP stream) is free, on that slot's stream, so requests that arrive while
others are on the GPU overlap with them (continuous batching: each dispatch launches every waiting request
whose catchment is idle); timesteps of one catchment stay in order because a catchment has at most one
request in flight. A request's latency is from its arrival to the completion of its timestep (HIP event on
its stream).

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
    states: list                      # two discharge buffers: states[current] holds the current discharge
    lateral: fw.PartitionedArray      # lateral inflow of the next timestep
    network: fw.FlowNetwork
    plans: list                       # plans[i]: timestep states[i] -> states[1 - i]
    resident_bytes: int
    current: int = 0
    busy: bool = False
    steps: int = 0

    @property
    def state(self):
        return self.states[self.current]


@dataclasses.dataclass(order=True)
class Request:
    arrival: float
    seq: int
    catchment: str = dataclasses.field(compare=False)
    seed: int = dataclasses.field(compare=False)
    # filled when it runs
    admitted: float | None = dataclasses.field(default=None, compare=False)
    done_event: torch.cuda.Event | None = dataclasses.field(default=None, compare=False)
    start_event: torch.cuda.Event | None = dataclasses.field(default=None, compare=False)
    result: fw.PartitionedArray | None = dataclasses.field(default=None, compare=False)
    latency: float | None = dataclasses.field(default=None, compare=False)


class AdmissionError(RuntimeError):
    pass


class Server:
    """Resident catchments and the request loop. beta, dt and dx are the scalars of every timestep.
    keep_results: keep every request's result (True), none (False), or those of the named catchments."""

    # resident bytes per cell: ldd 1, alpha 4, two states 8, lateral inflow 4, recorded schedule (at most 24),
    # the two plans' per-call buffers (at most 2 x 17)
    RESIDENT_BYTES_PER_CELL = 1 + 4 + 8 + 4 + 24 + 34

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
        self.completed: list[Request] = []
        self.cells_done = 0
        self._warm = False

    # -- resident catchments --------------------------------------------------------------------------
    def add_catchment(self, name, ldd, alpha, initial_state):
        """Make a catchment resident: record its network and capture its timestep. Raises AdmissionError
        beyond the HBM budget."""
        rows, cols = ldd.shape
        estimate = rows * cols * self.RESIDENT_BYTES_PER_CELL
        if self.resident_bytes + estimate > self.budget:
            raise AdmissionError(f"catchment {name}: {estimate} B exceed the HBM budget")
        network = fw.FlowNetwork(ldd, tile_shape=self.tile_shape)
        states = [fw.PartitionedArray(initial_state.data.clone(), initial_state.partition_shape),
                  fw.PartitionedArray(torch.zeros_like(initial_state.data), initial_state.partition_shape)]
        lateral = fw.PartitionedArray(torch.zeros_like(initial_state.data), initial_state.partition_shape)
        plans = [network.plan_kinematic_wave(states[i], lateral, alpha, self.beta, self.dt, self.dx, out=states[1 - i])
                 for i in range(2)]
        self.catchments[name] = Catchment(name, ldd, alpha, states, lateral, network, plans, estimate)
        self.resident_bytes += estimate
        self._warm = False
        return self.catchments[name]

    def remove_catchment(self, name):
        c = self.catchments[name]
        if c.busy:
            raise RuntimeError(f"catchment {name} has a request in flight")
        del self.catchments[name]
        for plan in c.plans:
            plan.close()
        c.network.close()
        self.resident_bytes -= c.resident_bytes

    def warm_up(self):
        """Launch every catchment's current timestep graph once on every slot's stream (the first launch of a
        graph on a stream carries one-time costs) into the catchment's next state buffer, which the next
        request overwrites: states are unchanged."""
        torch.cuda.synchronize(self.device)
        for stream in self.streams:
            with torch.cuda.stream(stream):
                for c in self.catchments.values():
                    c.plans[c.current].run()
            torch.cuda.synchronize(self.device)
        self._warm = True

    # -- one timestep ------------------------------------------------------------------------------------
    @staticmethod
    def lateral_inflow(seed, shape, device):
        return fw.PartitionedArray(syn.uniform_device(seed, syn.STREAM_LATERAL, 0.0, 1e-4, shape, torch.float32,
                                                      device))

    def _launch(self, request, slot):
        c = self.catchments[request.catchment]
        stream = self.streams[slot]
        keep = self.keep_results is True or (self.keep_results and request.catchment in self.keep_results)
        with torch.cuda.stream(stream):
            request.start_event = torch.cuda.Event(enable_timing=False)
            request.start_event.record(stream)
            syn.uniform_into(request.seed, syn.STREAM_LATERAL, 0.0, 1e-4, c.lateral.data)
            c.plans[c.current].run()
            c.current = 1 - c.current
            if keep:
                request.result = fw.PartitionedArray(c.state.data.clone(), c.state.partition_shape)
            request.done_event = torch.cuda.Event(enable_timing=False)
            request.done_event.record(stream)
        c.busy = True
        c.steps += 1

    def _finish(self, request, slot, now):
        c = self.catchments[request.catchment]
        c.busy = False
        request.latency = now - request.arrival
        self.free_slots.append(slot)
        self.cells_done += c.ldd.shape[0] * c.ldd.shape[1]
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
            # dispatch: FIFO over the waiting requests whose catchment is idle, while slots last (a request
            # needs no memory beyond its catchment's resident buffers)
            kept = collections.deque()
            while waiting:
                request = waiting.popleft()
                c = self.catchments.get(request.catchment)
                if c is None:
                    raise AdmissionError(f"request for {request.catchment}: no such resident catchment")
                if c.busy or not self.free_slots:
                    kept.append(request)
                    continue
                request.admitted = clock() - t0
                slot = self.free_slots.pop()
                self._launch(request, slot)
                in_flight.append((request, slot))
            waiting = kept
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
