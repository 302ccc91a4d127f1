"""Python front end of the routing path, mirroring lue.framework's routing operations.

Reference bindings: source/framework/python/source/algorithm/routing_operation/{accu,accu_threshold,
accu_fraction,accu_capacity,accu_trigger,accu_info,kinematic_wave,inflow_count,d8_flow_direction,upstream,
downstream}.cpp, which call lue::value_policies::* on PartitionedArray.

PartitionedArray here is device resident: one contiguous row-major raster in HBM (a torch tensor used as
device memory) plus LUE's partition shape, which is kept as metadata: only accu_info's classes depend on
it. A Scalar argument is a Python number or numpy scalar (LUE's Scalar<E> and plain-value overloads).

Operations are asynchronous on torch's current stream, like LUE's operations, which return future-backed
partitions at once (partitioned_array_decl.hpp:23-98): they enqueue their kernels and return; reading a
result (numpy(), torch ops) on the same stream is ordered after them. The routing operations run on the
flow network (recorded schedule) of their ldd, which a PartitionedArray of uint8 keeps after its first
routing operation, until its data change (torch's version counter of the tensor) or it goes away.
"""
from __future__ import annotations

import numpy as np
import torch

from . import _lib

FLOAT_TYPES = {torch.float32: "f32", torch.float64: "f64"}


def _default_device():
    if not torch.cuda.is_available():
        raise RuntimeError("lue_amd needs a GPU: no HIP device is visible")
    return torch.device("cuda", torch.cuda.current_device())


class PartitionedArray:
    """Device-resident 2-D array with LUE's partitioning metadata.

    Mirrors lue::PartitionedArray<E, 2> (reference source/framework/partitioned_array/include/lue/framework/
    partitioned_array_decl.hpp:23-98): shape, partition shape, shape in partitions (the last partition
    along each dimension absorbs the remainder, domain_decomposition.hpp:91-183).
    """

    def __init__(self, data: torch.Tensor, partition_shape=None):
        if data.dim() != 2:
            raise ValueError("a PartitionedArray is 2-D")
        if data.device.type != "cuda":
            raise ValueError("a PartitionedArray lives in device memory")
        self.data = data.contiguous()
        shape = tuple(self.data.shape)
        if partition_shape is None:
            partition_shape = shape
        partition_shape = tuple(int(v) for v in partition_shape)
        if any(p <= 0 or p > s for p, s in zip(partition_shape, shape)):
            raise ValueError(f"Partition shape {partition_shape} must fit in array shape {shape}")
        self.partition_shape = partition_shape

    @property
    def shape(self):
        return tuple(self.data.shape)

    @property
    def dtype(self):
        return self.data.dtype

    @property
    def shape_in_partitions(self):
        return tuple(s // p for s, p in zip(self.shape, self.partition_shape))

    @property
    def nr_partitions(self):
        a, b = self.shape_in_partitions
        return a * b

    def numpy(self):
        return self.data.cpu().numpy()

    def _routing_network(self, operation):
        """The flow network of this ldd array for the per-call routing operation `operation` (its tile shape,
        PER_CALL_TILE_SHAPES): recorded at the first call (which waits for the recording), kept while the
        tensor and its contents are unchanged, so later calls only enqueue kernels; one per tile shape. None
        when the per-call schedules are not kept (set_network_cache(0))."""
        if _lib.network_cache_entries() == 0:
            return None
        shape = PER_CALL_TILE_SHAPES.get(operation, PER_CALL_TILE_SHAPES["accu"])
        key = (self.data.data_ptr(), tuple(self.data.shape), self.data._version, _lib.sweep_schedule(),
               self.data.device)
        networks = self.__dict__.get("_networks")
        if networks is None or networks[0] != key:
            if networks is not None:
                for net in networks[1].values():
                    net.close()
            networks = (key, {})
            self._networks = networks
        net = networks[1].get(shape)
        if net is None:
            net = FlowNetwork(self, tile_shape=shape, schedule=_lib.sweep_schedule())
            networks[1][shape] = net
        return net

    @classmethod
    def from_numpy(cls, array, partition_shape=None, device=None):
        device = device or _default_device()
        return cls(torch.from_numpy(np.ascontiguousarray(array)).to(device), partition_shape)


def create_partitioned_array(array_shape, partition_shape, dtype, fill_value, device=None):
    """lue.framework.create_array / lue::create_partitioned_array(shape, partition_shape, fill)."""
    device = device or _default_device()
    data = torch.full(tuple(array_shape), fill_value, dtype=dtype, device=device)
    return PartitionedArray(data, partition_shape)


def _stream():
    return ctypes_stream(torch.cuda.current_stream())


def ctypes_stream(stream):
    return stream.cuda_stream


def _ptr(t: torch.Tensor):
    return t.data_ptr()


def _verify_compatible(flow_direction, *arrays):
    """detail::verify_compatible (ALG/detail/verify_compatible.hpp:22-41): same shape, same partitioning."""
    for a in arrays:
        if isinstance(a, PartitionedArray):
            if a.shape != flow_direction.shape:
                raise ValueError(f"Shapes differ: {a.shape} != {flow_direction.shape}")
            if a.shape_in_partitions != flow_direction.shape_in_partitions:
                raise ValueError(f"Shapes in partitions differ: {a.shape_in_partitions} != "
                                 f"{flow_direction.shape_in_partitions}")
            if a.data.device != flow_direction.data.device:
                raise ValueError("arrays live on different devices")


def _element_type(*args):
    types = {a.dtype for a in args if isinstance(a, PartitionedArray)}
    if len(types) > 1:
        raise TypeError(f"material arguments must share one element type, got {types}")
    dtype = types.pop() if types else torch.float32
    if dtype not in FLOAT_TYPES:
        raise TypeError(f"material element type must be float32 or float64, not {dtype}")
    return dtype


def _scalar(value, dtype):
    """A raster as it is; a Scalar as a float (the network entry points take it in the operation's type)."""
    return value if isinstance(value, PartitionedArray) else float(value)


def _arg(value, dtype):
    """(pointer or None, scalar) for a raster or Scalar argument."""
    if isinstance(value, PartitionedArray):
        if value.dtype != dtype:
            raise TypeError(f"expected {dtype}, got {value.dtype}")
        return _ptr(value.data), 0.0
    return None, float(value)


def _check_ldd(flow_direction):
    if not isinstance(flow_direction, PartitionedArray) or flow_direction.dtype != torch.uint8:
        raise TypeError("flow_direction must be a PartitionedArray of uint8")


def _output(flow_direction, dtype):
    return PartitionedArray(torch.empty(flow_direction.shape, dtype=dtype, device=flow_direction.data.device),
                            flow_direction.partition_shape)


def inflow_count(flow_direction):
    """lue::inflow_count<uint8>(flow_direction) (ALG/inflow_count.hpp:29-42)."""
    _check_ldd(flow_direction)
    lib = _lib.load()
    out = _output(flow_direction, torch.uint8)
    rows, cols = flow_direction.shape
    _lib.check(lib.lue_amd_inflow_count_u8(_ptr(flow_direction.data), rows, cols, _ptr(out.data), _stream()))
    return out


# Cell classes of accu_info (ALG/accu_info.hpp:27-63); no-data is 255.
ridge_cell = 11
intra_partition_stream_cell = 12
partition_output_cell = 13
partition_input_cell = 14
partition_input_output_cell = 15
inter_partition_stream_cell = 16
sink_cell = 17
confluence_cell = 18


def accu_info(flow_direction):
    """lue.framework.accu_info(flow_direction) (routing_operation/accu_info.cpp:10-17,
    ALG/value_policies/accu_info.hpp:19-28): the class the accumulating router leaves in each cell for the
    array's partition shape, uint8. A partition input cell that also has an upstream neighbour in its own
    partition waiting for another partition is a partition input cell here; in LUE it can also end up as
    the class of a cell without inflow from other partitions, depending on which arrives last."""
    _check_ldd(flow_direction)
    lib = _lib.load()
    out = _output(flow_direction, torch.uint8)
    rows, cols = flow_direction.shape
    prows, pcols = flow_direction.partition_shape
    _lib.check(lib.lue_amd_accu_info_u8(_ptr(flow_direction.data), rows, cols, prows, pcols, _ptr(out.data),
                                        _stream()))
    return out


def d8_flow_direction(elevation):
    """lue.framework.d8_flow_direction(elevation) -> uint8 ldd with the elevation's partitioning
    (value_policies/d8_flow_direction.hpp:46-55, definition/d8_flow_direction.hpp:12-173). NaN elevation is
    no-data."""
    if not isinstance(elevation, PartitionedArray):
        raise TypeError("elevation must be a PartitionedArray")
    dtype = elevation.data.dtype
    if dtype not in FLOAT_TYPES:
        raise TypeError(f"unsupported elevation element type {dtype}")
    lib = _lib.load()
    out = _output(elevation, torch.uint8)
    rows, cols = elevation.shape
    fn = getattr(lib, f"lue_amd_d8_flow_direction_{FLOAT_TYPES[dtype]}")
    _lib.check(fn(_ptr(elevation.data), rows, cols, _ptr(out.data), _stream()))
    return out


# material element types of LUE's default build (environment/cmake/LueConfiguration.cmake:171-228)
MATERIAL_TYPES = {torch.uint8: "u8", torch.uint32: "u32", torch.uint64: "u64", torch.int8: "i8",
                  torch.int32: "i32", torch.int64: "i64", torch.float32: "f32", torch.float64: "f64"}


def _flow_neighbour(op, flow_direction, material):
    _check_ldd(flow_direction)
    if not isinstance(material, PartitionedArray):
        raise TypeError("material must be a PartitionedArray")
    _verify_compatible(flow_direction, material)
    if material.dtype not in MATERIAL_TYPES:
        raise TypeError(f"unsupported material element type {material.dtype}")
    lib = _lib.load()
    out = _output(flow_direction, material.dtype)
    rows, cols = flow_direction.shape
    fn = getattr(lib, f"lue_amd_{op}_{MATERIAL_TYPES[material.dtype]}")
    _lib.check(fn(_ptr(flow_direction.data), _ptr(material.data), rows, cols, _ptr(out.data), _stream()))
    return out


def upstream(flow_direction, material):
    """lue.framework.upstream(flow_direction, material): per cell, the sum of the material of the neighbours
    that drain into it (value_policies/upstream.hpp:16-30, definition/upstream.hpp:36-190)."""
    return _flow_neighbour("upstream", flow_direction, material)


def downstream(flow_direction, material):
    """lue.framework.downstream(flow_direction, material): per cell, the material of its downstream
    neighbour (value_policies/downstream.hpp:16-30, definition/downstream.hpp:36-186)."""
    return _flow_neighbour("downstream", flow_direction, material)


def downstream_distance(flow_direction, cell_size, dtype=np.float32):
    """lue.framework.downstream_distance(flow_direction, cell_size, dtype): per cell, the distance to its
    downstream cell (python/source/algorithm/routing_operation/downstream_distance.cpp:14-23, value policies:
    a no-data ldd gives NaN; definition/downstream_distance.hpp:13-75). dtype selects float32 or float64 by
    its item size, as the binding does."""
    _check_ldd(flow_direction)
    size = np.dtype(dtype).itemsize
    if size not in (4, 8):
        raise TypeError(f"unsupported distance element type {dtype}")
    torch_dtype = torch.float32 if size == 4 else torch.float64
    lib = _lib.load()
    out = _output(flow_direction, torch_dtype)
    rows, cols = flow_direction.shape
    fn = getattr(lib, f"lue_amd_downstream_distance_{FLOAT_TYPES[torch_dtype]}")
    _lib.check(fn(_ptr(flow_direction.data), rows, cols, float(cell_size), 1, _ptr(out.data), _stream()))
    return out


def accu(flow_direction, inflow):
    """lue.framework.accu(flow_direction, inflow) -> outflow (value_policies/accu.hpp:33-55)."""
    _check_ldd(flow_direction)
    _verify_compatible(flow_direction, inflow)
    dtype = _element_type(inflow)
    net = flow_direction._routing_network("accu")
    if net is not None:
        return net.accu(inflow if isinstance(inflow, PartitionedArray) else _scalar(inflow, dtype), dtype=dtype)
    lib = _lib.load()
    out = _output(flow_direction, dtype)
    p, s = _arg(inflow, dtype)
    rows, cols = flow_direction.shape
    fn = getattr(lib, f"lue_amd_accu_{FLOAT_TYPES[dtype]}")
    _lib.check(fn(_ptr(flow_direction.data), rows, cols, p, s, _ptr(out.data), _stream()))
    return out


def _split(name, flow_direction, inflow, param):
    _check_ldd(flow_direction)
    _verify_compatible(flow_direction, inflow, param)
    dtype = _element_type(inflow, param)
    net = flow_direction._routing_network(name)
    if net is not None:
        return net._split(name[len("accu_"):], _scalar(inflow, dtype), _scalar(param, dtype), dtype=dtype)
    lib = _lib.load()
    out = _output(flow_direction, dtype)
    rem = _output(flow_direction, dtype)
    p1, s1 = _arg(inflow, dtype)
    p2, s2 = _arg(param, dtype)
    rows, cols = flow_direction.shape
    fn = getattr(lib, f"lue_amd_{name}_{FLOAT_TYPES[dtype]}")
    _lib.check(fn(_ptr(flow_direction.data), rows, cols, p1, s1, p2, s2, _ptr(out.data), _ptr(rem.data),
                  _stream()))
    return out, rem


def accu_threshold(flow_direction, inflow, threshold):
    """lue.framework.accu_threshold -> (outflow, remainder) (value_policies/accu_threshold.hpp:34-87)."""
    return _split("accu_threshold", flow_direction, inflow, threshold)


def accu_fraction(flow_direction, inflow, fraction):
    """lue.framework.accu_fraction -> (outflow, remainder) (value_policies/accu_fraction.hpp)."""
    return _split("accu_fraction", flow_direction, inflow, fraction)


def accu_capacity(flow_direction, inflow, capacity):
    """lue.framework.accu_capacity -> (outflow, remainder) (value_policies/accu_capacity.hpp:34-88)."""
    return _split("accu_capacity", flow_direction, inflow, capacity)


def accu_trigger(flow_direction, inflow, trigger):
    """lue.framework.accu_trigger -> (outflow, remainder) (value_policies/accu_trigger.hpp:34-88)."""
    return _split("accu_trigger", flow_direction, inflow, trigger)


def kinematic_wave(flow_direction, current_outflow, inflow, alpha, beta, time_step_duration, channel_length):
    """lue.framework.kinematic_wave -> new outflow (value_policies/kinematic_wave.hpp:43-91)."""
    _check_ldd(flow_direction)
    if not isinstance(current_outflow, PartitionedArray) or not isinstance(inflow, PartitionedArray):
        raise TypeError("current_outflow and inflow must be PartitionedArrays")
    _verify_compatible(flow_direction, current_outflow, inflow, alpha, beta, channel_length)
    dtype = _element_type(current_outflow, inflow, alpha, beta, channel_length)
    if isinstance(time_step_duration, PartitionedArray):
        raise TypeError("time_step_duration is a Scalar")
    net = flow_direction._routing_network("kinematic_wave")
    if net is not None:
        return net.kinematic_wave(current_outflow, inflow, _scalar(alpha, dtype), _scalar(beta, dtype),
                                  float(time_step_duration), _scalar(channel_length, dtype), dtype=dtype)
    lib = _lib.load()
    out = _output(flow_direction, dtype)
    pa, sa = _arg(alpha, dtype)
    pb, sb = _arg(beta, dtype)
    pl, sl = _arg(channel_length, dtype)
    rows, cols = flow_direction.shape
    fn = getattr(lib, f"lue_amd_kinematic_wave_{FLOAT_TYPES[dtype]}")
    _lib.check(fn(_ptr(flow_direction.data), rows, cols, _ptr(current_outflow.data), _ptr(inflow.data), pa, sa,
                  pb, sb, float(time_step_duration), pl, sl, _ptr(out.data), _stream()))
    return out


def iterate_to_new_discharge(upstream, current, lateral, alpha, beta, dt, dx):
    """Elementwise lue::detail::iterate_to_new_discharge over device tensors of one float type."""
    tensors = [upstream, current, lateral, alpha, beta, dt, dx]
    dtype = upstream.dtype
    tensors = [t.to(dtype).contiguous() for t in tensors]
    lib = _lib.load()
    out = torch.empty_like(tensors[0])
    fn = getattr(lib, f"lue_amd_iterate_to_new_discharge_{FLOAT_TYPES[dtype]}")
    _lib.check(fn(*[_ptr(t) for t in tensors], _ptr(out), tensors[0].numel(), _stream()))
    return out


# ---------------------------------------------------------------------------------------------------
# Flow networks: resident ldd rasters with their recorded sweep schedule (C ABI lue_amd_network_*)

TILE_SHAPES = {(32, 32): 0, (64, 32): 1, (64, 64): 2, (32, 64): 3}

# Tile shape (width, height) of the per-call operations, the same as the C ABI's per-call entry points: the
# kinematic wave needs the occupancy of the smallest tiles, the accu family gains from larger ones (fewer
# cells left to the chain phase).
PER_CALL_TILE_SHAPES = {"accu": (64, 32), "kinematic_wave": (32, 32)}


def _pointer_array(values):
    import ctypes

    arr = (ctypes.c_void_p * len(values))(*values)
    return arr


class FlowNetwork:
    """One or more catchments (uint8 ldd PartitionedArrays) and the schedule of their routing sweep.

    The schedule depends on the ldd only, so it is recorded once; every operation then runs as a fixed
    launch sequence over all catchments at once. The ldd arrays must not change while the network lives.
    tile_shape: (width, height) of the sweep tiles, one of TILE_SHAPES; schedule: "chains" (the cells the first
    round of tiles leaves pending are solved along heavy-path chains) or "rounds" (further rounds of tiles).
    """

    def __init__(self, flow_directions, tile_shape=(32, 32), schedule="chains"):
        import ctypes

        if isinstance(flow_directions, PartitionedArray):
            flow_directions = [flow_directions]
        for ldd in flow_directions:
            _check_ldd(ldd)
        # views of the same tensors: the network keeps the ldd data alive (it borrows the device pointers)
        # without referring to the arrays themselves, which may hold this network
        self.flow_directions = [PartitionedArray(ldd.data, ldd.partition_shape) for ldd in flow_directions]
        lib = _lib.load()
        n = len(self.flow_directions)
        ptrs = _pointer_array([_ptr(a.data) for a in self.flow_directions])
        rows = (ctypes.c_int64 * n)(*[a.shape[0] for a in self.flow_directions])
        cols = (ctypes.c_int64 * n)(*[a.shape[1] for a in self.flow_directions])
        handle = ctypes.c_void_p()
        self._handle = None
        _lib.check(lib.lue_amd_network_create(n, ptrs, rows, cols, TILE_SHAPES[tuple(tile_shape)],
                                              _lib.SCHEDULES[schedule],
                                              _stream(), ctypes.byref(handle)))
        self._handle = handle
        self.tile_shape = tuple(tile_shape)

    def close(self):
        if getattr(self, "_handle", None):
            _lib.load().lue_amd_network_destroy(self._handle)
            self._handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __len__(self):
        return len(self.flow_directions)

    def info(self):
        import ctypes

        keys = ["catchments", "cells", "tiles", "rounds", "pending_cells", "chains", "levels", "no_data_cells",
                "tile_shape", "schedule", "longest_chain", "side_refs", "tile_blocks", "tile_critical_blocks",
                "wide_threshold", "tile_queue", "lane_limit"]
        values = (ctypes.c_int64 * len(keys))()
        _lib.check(_lib.load().lue_amd_network_info(self._handle, values, len(keys)))
        return dict(zip(keys, list(values)))

    def set_wide_threshold(self, positions):
        """Chain levels whose longest work item has at least `positions` positions run one work item per
        workgroup of 16 wavefronts (0: none); the results are the same either way. Applies to the
        operations and plans made from then on."""
        _lib.check(_lib.load().lue_amd_network_set_wide_threshold(self._handle, int(positions)))

    def set_lane_limit(self, positions):
        """Work items of a chain level with at most `positions` positions are walked one per lane, cell after
        cell, the longer ones one per wavefront or workgroup (0, the default: none). The results agree with
        the other schedule to the operations' tolerances; whether an item is walked depends on its length
        only, not on the other catchments of a batch. Applies to the operations and plans made from then on."""
        _lib.check(_lib.load().lue_amd_network_set_lane_limit(self._handle, int(positions)))

    def set_tile_queue(self, on):
        """Launch the first round as a persistent grid fed by a tile work queue (True) or one workgroup per
        tile (False, the default); the results are the same. Applies to the operations and plans made from
        then on."""
        _lib.check(_lib.load().lue_amd_network_set_tile_queue(self._handle, 1 if on else 0))

    def level_info(self):
        """Per chain level: positions, work items, positions of the longest work item, side references."""
        import ctypes

        keys = ["positions", "items", "longest_item", "side_refs"]
        values = (ctypes.c_int64 * len(keys))()
        levels = []
        for level in range(self.info()["levels"]):
            _lib.check(_lib.load().lue_amd_network_level_info(self._handle, level, values, len(keys)))
            levels.append(dict(zip(keys, list(values))))
        return levels

    # -- argument helpers -------------------------------------------------------------------------
    def _rasters(self, value, dtype, name):
        """Per-catchment pointer array (or None and a scalar)."""
        if isinstance(value, (list, tuple)):
            if len(value) != len(self):
                raise ValueError(f"{name}: one raster per catchment")
            for v, ldd in zip(value, self.flow_directions):
                if not isinstance(v, PartitionedArray) or v.dtype != dtype:
                    raise TypeError(f"{name}: PartitionedArrays of {dtype} expected")
                _verify_compatible(ldd, v)
            return _pointer_array([_ptr(v.data) for v in value]), 0.0, value
        if isinstance(value, PartitionedArray):
            if len(self) != 1:
                raise ValueError(f"{name}: one raster per catchment")
            return self._rasters([value], dtype, name)
        return None, float(value), None

    def _outputs(self, dtype, out=None):
        """Per-catchment output pointer array: new rasters, or the caller's, which must match each
        catchment's ldd in shape and device and have the operation's element type."""
        if out is None:
            out = [_output(ldd, dtype) for ldd in self.flow_directions]
        else:
            if isinstance(out, PartitionedArray):
                out = [out]
            out = list(out)
            if len(out) != len(self):
                raise ValueError(f"one output raster per catchment: {len(out)} given for {len(self)}")
            for o, ldd in zip(out, self.flow_directions):
                if not isinstance(o, PartitionedArray):
                    raise TypeError("outputs must be PartitionedArrays")
                if o.shape != ldd.shape:
                    raise ValueError(f"Shapes differ: {o.shape} != {ldd.shape}")
                if o.dtype != dtype:
                    raise TypeError(f"output element type {o.dtype} differs from the operation's {dtype}")
                if o.data.device != ldd.data.device:
                    raise ValueError("output lives on another device than the ldd")
        return _pointer_array([_ptr(o.data) for o in out]), out

    @staticmethod
    def _dtype(*values, dtype=None):
        if dtype is not None:
            return dtype
        for v in values:
            if isinstance(v, (list, tuple)) and v:
                return v[0].dtype
            if isinstance(v, PartitionedArray):
                return v.dtype
        return torch.float32

    def _unwrap(self, outs):
        return outs[0] if len(self) == 1 else outs

    # -- plans: one operation with fixed rasters, captured into a HIP graph ---------------------------
    def _plan(self, name, fn_args, outs, keep):
        import ctypes

        handle = ctypes.c_void_p()
        _lib.check(getattr(_lib.load(), name)(self._handle, *fn_args, ctypes.byref(handle)))
        return Plan(self, handle, outs, keep)

    def plan_accu(self, inflow, out=None, dtype=None):
        """accu with these rasters as a Plan (one graph launch per run); outputs in plan.outputs."""
        dtype = self._dtype(inflow, dtype=dtype)
        p, s, keep = self._rasters(inflow, dtype, "inflow")
        po, outs = self._outputs(dtype, out)
        return self._plan(f"lue_amd_plan_accu_{FLOAT_TYPES[dtype]}", (p, s, po), [self._unwrap(outs)], (keep, outs))

    def _plan_split(self, name, inflow, param, out=None, remainder=None, dtype=None):
        dtype = self._dtype(inflow, param, dtype=dtype)
        p1, s1, k1 = self._rasters(inflow, dtype, "inflow")
        p2, s2, k2 = self._rasters(param, dtype, name)
        po, outs = self._outputs(dtype, out)
        pr, rems = self._outputs(dtype, remainder)
        return self._plan(f"lue_amd_plan_accu_{name}_{FLOAT_TYPES[dtype]}", (p1, s1, p2, s2, po, pr),
                          [self._unwrap(outs), self._unwrap(rems)], (k1, k2, outs, rems))

    def plan_accu_threshold(self, inflow, threshold, out=None, remainder=None, dtype=None):
        return self._plan_split("threshold", inflow, threshold, out, remainder, dtype)

    def plan_accu_fraction(self, inflow, fraction, out=None, remainder=None, dtype=None):
        return self._plan_split("fraction", inflow, fraction, out, remainder, dtype)

    def plan_accu_capacity(self, inflow, capacity, out=None, remainder=None, dtype=None):
        return self._plan_split("capacity", inflow, capacity, out, remainder, dtype)

    def plan_accu_trigger(self, inflow, trigger, out=None, remainder=None, dtype=None):
        return self._plan_split("trigger", inflow, trigger, out, remainder, dtype)

    def plan_kinematic_wave(self, current_outflow, inflow, alpha, beta, time_step_duration, channel_length,
                            out=None, dtype=None):
        dtype = self._dtype(current_outflow, inflow, dtype=dtype)
        pc, _, kc = self._rasters(current_outflow, dtype, "current_outflow")
        pq, _, kq = self._rasters(inflow, dtype, "inflow")
        if pc is None or pq is None:
            raise TypeError("current_outflow and inflow are rasters")
        pa, sa, ka = self._rasters(alpha, dtype, "alpha")
        pb, sb, kb = self._rasters(beta, dtype, "beta")
        pl, sl, kl = self._rasters(channel_length, dtype, "channel_length")
        po, outs = self._outputs(dtype, out)
        return self._plan(f"lue_amd_plan_kinematic_wave_{FLOAT_TYPES[dtype]}",
                          (pc, pq, pa, sa, pb, sb, float(time_step_duration), pl, sl, po), [self._unwrap(outs)],
                          (kc, kq, ka, kb, kl, outs))

    # -- operations (asynchronous on the current stream) --------------------------------------------
    def accu(self, inflow, out=None, dtype=None):
        dtype = self._dtype(inflow, dtype=dtype)
        p, s, keep = self._rasters(inflow, dtype, "inflow")
        po, outs = self._outputs(dtype, out)
        fn = getattr(_lib.load(), f"lue_amd_network_accu_{FLOAT_TYPES[dtype]}")
        _lib.check(fn(self._handle, p, s, po, _stream()))
        del keep
        return self._unwrap(outs)

    def _split(self, name, inflow, param, out=None, remainder=None, dtype=None):
        dtype = self._dtype(inflow, param, dtype=dtype)
        p1, s1, k1 = self._rasters(inflow, dtype, "inflow")
        p2, s2, k2 = self._rasters(param, dtype, name)
        po, outs = self._outputs(dtype, out)
        pr, rems = self._outputs(dtype, remainder)
        fn = getattr(_lib.load(), f"lue_amd_network_accu_{name}_{FLOAT_TYPES[dtype]}")
        _lib.check(fn(self._handle, p1, s1, p2, s2, po, pr, _stream()))
        del k1, k2
        return self._unwrap(outs), self._unwrap(rems)

    def accu_threshold(self, inflow, threshold, out=None, remainder=None):
        return self._split("threshold", inflow, threshold, out, remainder)

    def accu_fraction(self, inflow, fraction, out=None, remainder=None):
        return self._split("fraction", inflow, fraction, out, remainder)

    def accu_capacity(self, inflow, capacity, out=None, remainder=None):
        return self._split("capacity", inflow, capacity, out, remainder)

    def accu_trigger(self, inflow, trigger, out=None, remainder=None):
        return self._split("trigger", inflow, trigger, out, remainder)

    def kinematic_wave(self, current_outflow, inflow, alpha, beta, time_step_duration, channel_length, out=None,
                       dtype=None):
        dtype = self._dtype(current_outflow, inflow, dtype=dtype)
        pc, _, kc = self._rasters(current_outflow, dtype, "current_outflow")
        pq, _, kq = self._rasters(inflow, dtype, "inflow")
        if pc is None or pq is None:
            raise TypeError("current_outflow and inflow are rasters")
        pa, sa, ka = self._rasters(alpha, dtype, "alpha")
        pb, sb, kb = self._rasters(beta, dtype, "beta")
        pl, sl, kl = self._rasters(channel_length, dtype, "channel_length")
        po, outs = self._outputs(dtype, out)
        fn = getattr(_lib.load(), f"lue_amd_network_kinematic_wave_{FLOAT_TYPES[dtype]}")
        _lib.check(fn(self._handle, pc, pq, pa, sa, pb, sb, float(time_step_duration), pl, sl, po, _stream()))
        del kc, kq, ka, kb, kl
        return self._unwrap(outs)


class Plan:
    """One operation of a FlowNetwork with fixed rasters, its launch sequence captured into a HIP graph
    (C ABI lue_amd_plan_*): run() replays it as one graph launch on the current stream. The rasters' contents
    may change between runs (e.g. a new lateral inflow), not the rasters themselves; the plan keeps them and
    its network alive. outputs: the output raster(s), as the network operation returns them."""

    def __init__(self, network, handle, outputs, keep):
        self.network = network
        self._handle = handle
        self.outputs = outputs[0] if len(outputs) == 1 else tuple(outputs)
        self._keep = keep

    def run(self):
        _lib.check(_lib.load().lue_amd_plan_launch(self._handle, _stream()))
        return self.outputs

    def info(self):
        import ctypes

        values = (ctypes.c_int64 * 2)()
        _lib.check(_lib.load().lue_amd_plan_info(self._handle, values, 2))
        return {"launches": int(values[0]), "buffer_bytes": int(values[1])}

    def close(self):
        if getattr(self, "_handle", None):
            _lib.load().lue_amd_plan_destroy(self._handle)
            self._handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
