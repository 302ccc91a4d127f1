"""Loader of liblue_amd.so, the HIP implementation behind the C ABI in include/lue_amd/routing.h.

The library is built in-tree (lue_amd/build.py). There is no fallback: if the library is missing or a
call fails, an exception is raised.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liblue_amd.so")

_lib = None

c_p = ctypes.c_void_p
c_i64 = ctypes.c_int64
c_u64 = ctypes.c_uint64
c_f = ctypes.c_float
c_d = ctypes.c_double
c_int = ctypes.c_int
c_size = ctypes.c_size_t

# name -> argument types (restype int unless listed in _RESTYPES)
SIGNATURES = {
    "lue_amd_status_string": [c_int],
    "lue_amd_last_error": [],
    "lue_amd_version": [],
    "lue_amd_set_sweep_options": [c_int],
    "lue_amd_set_network_cache": [c_int],
    "lue_amd_trim_scratch": [],
    "lue_amd_network_create": [c_int, c_p, c_p, c_p, c_int, c_int, c_p, ctypes.POINTER(c_p)],
    "lue_amd_network_destroy": [c_p],
    "lue_amd_network_info": [c_p, c_p, c_int],
    "lue_amd_network_level_info": [c_p, c_int, c_p, c_int],
    "lue_amd_network_set_wide_threshold": [c_p, c_i64],
    "lue_amd_network_set_tile_queue": [c_p, c_int],
    "lue_amd_network_set_lane_limit": [c_p, c_i64],
    "lue_amd_network_accu_f32": [c_p, c_p, c_f, c_p, c_p],
    "lue_amd_network_accu_f64": [c_p, c_p, c_d, c_p, c_p],
    "lue_amd_network_accu_threshold_f32": [c_p, c_p, c_f, c_p, c_f, c_p, c_p, c_p],
    "lue_amd_network_accu_threshold_f64": [c_p, c_p, c_d, c_p, c_d, c_p, c_p, c_p],
    "lue_amd_network_accu_fraction_f32": [c_p, c_p, c_f, c_p, c_f, c_p, c_p, c_p],
    "lue_amd_network_accu_fraction_f64": [c_p, c_p, c_d, c_p, c_d, c_p, c_p, c_p],
    "lue_amd_network_accu_capacity_f32": [c_p, c_p, c_f, c_p, c_f, c_p, c_p, c_p],
    "lue_amd_network_accu_capacity_f64": [c_p, c_p, c_d, c_p, c_d, c_p, c_p, c_p],
    "lue_amd_network_accu_trigger_f32": [c_p, c_p, c_f, c_p, c_f, c_p, c_p, c_p],
    "lue_amd_network_accu_trigger_f64": [c_p, c_p, c_d, c_p, c_d, c_p, c_p, c_p],
    "lue_amd_network_kinematic_wave_f32": [c_p, c_p, c_p, c_p, c_f, c_p, c_f, c_f, c_p, c_f, c_p, c_p],
    "lue_amd_network_kinematic_wave_f64": [c_p, c_p, c_p, c_p, c_d, c_p, c_d, c_d, c_p, c_d, c_p, c_p],
    "lue_amd_network_kinematic_wave_f32_noalloc": [c_p, c_p, c_p, c_p, c_f, c_p, c_f, c_f, c_p, c_f, c_p, c_p],
    "lue_amd_plan_launch": [c_p, c_p],
    "lue_amd_plan_destroy": [c_p],
    "lue_amd_plan_info": [c_p, c_p, c_int],
    "lue_amd_plan_accu_f32": [c_p, c_p, c_f, c_p, ctypes.POINTER(c_p)],
    "lue_amd_plan_accu_f64": [c_p, c_p, c_d, c_p, ctypes.POINTER(c_p)],
    "lue_amd_plan_kinematic_wave_f32": [c_p, c_p, c_p, c_p, c_f, c_p, c_f, c_f, c_p, c_f, c_p, ctypes.POINTER(c_p)],
    "lue_amd_plan_kinematic_wave_f64": [c_p, c_p, c_p, c_p, c_d, c_p, c_d, c_d, c_p, c_d, c_p, ctypes.POINTER(c_p)],
    "lue_amd_profile_enable": [c_int],
    "lue_amd_profile_read": [ctypes.POINTER(c_d), ctypes.POINTER(c_d)] + [ctypes.POINTER(ctypes.c_longlong)] * 3,
    "lue_amd_profile_read_op": [c_int, c_int] + [ctypes.POINTER(c_d)] * 4 + [ctypes.POINTER(ctypes.c_longlong)],
    "lue_amd_device_malloc": [ctypes.POINTER(c_p), c_size],
    "lue_amd_device_free": [c_p],
    "lue_amd_copy_to_device": [c_p, c_p, c_size, c_p],
    "lue_amd_copy_to_host": [c_p, c_p, c_size, c_p],
    "lue_amd_stream_synchronize": [c_p],
    "lue_amd_event_create": [ctypes.POINTER(c_p)],
    "lue_amd_event_destroy": [c_p],
    "lue_amd_event_record": [c_p, c_p],
    "lue_amd_stream_wait_event": [c_p, c_p],
    "lue_amd_event_synchronize": [c_p],
    "lue_amd_event_query": [c_p, ctypes.POINTER(c_int)],
    "lue_amd_inflow_count_u8": [c_p, c_i64, c_i64, c_p, c_p],
    "lue_amd_accu_info_u8": [c_p, c_i64, c_i64, c_i64, c_i64, c_p, c_p],
    "lue_amd_d8_flow_direction_f32": [c_p, c_i64, c_i64, c_p, c_p],
    "lue_amd_d8_flow_direction_f64": [c_p, c_i64, c_i64, c_p, c_p],
    "lue_amd_accu_f32": [c_p, c_i64, c_i64, c_p, c_f, c_p, c_p],
    "lue_amd_accu_f64": [c_p, c_i64, c_i64, c_p, c_d, c_p, c_p],
    "lue_amd_accu_threshold_f32": [c_p, c_i64, c_i64, c_p, c_f, c_p, c_f, c_p, c_p, c_p],
    "lue_amd_accu_threshold_f64": [c_p, c_i64, c_i64, c_p, c_d, c_p, c_d, c_p, c_p, c_p],
    "lue_amd_accu_fraction_f32": [c_p, c_i64, c_i64, c_p, c_f, c_p, c_f, c_p, c_p, c_p],
    "lue_amd_accu_fraction_f64": [c_p, c_i64, c_i64, c_p, c_d, c_p, c_d, c_p, c_p, c_p],
    "lue_amd_accu_capacity_f32": [c_p, c_i64, c_i64, c_p, c_f, c_p, c_f, c_p, c_p, c_p],
    "lue_amd_accu_capacity_f64": [c_p, c_i64, c_i64, c_p, c_d, c_p, c_d, c_p, c_p, c_p],
    "lue_amd_accu_trigger_f32": [c_p, c_i64, c_i64, c_p, c_f, c_p, c_f, c_p, c_p, c_p],
    "lue_amd_accu_trigger_f64": [c_p, c_i64, c_i64, c_p, c_d, c_p, c_d, c_p, c_p, c_p],
    "lue_amd_kinematic_wave_f32": [c_p, c_i64, c_i64, c_p, c_p, c_p, c_f, c_p, c_f, c_f, c_p, c_f, c_p, c_p],
    "lue_amd_kinematic_wave_f64": [c_p, c_i64, c_i64, c_p, c_p, c_p, c_d, c_p, c_d, c_d, c_p, c_d, c_p, c_p],
    "lue_amd_iterate_to_new_discharge_f32": [c_p] * 8 + [c_i64, c_p],
    "lue_amd_iterate_to_new_discharge_f64": [c_p] * 8 + [c_i64, c_p],
    "lue_amd_synthetic_ldd": [c_u64, c_i64, c_i64, c_p, c_p],
    "lue_amd_synthetic_dem_f32": [c_u64, c_i64, c_i64, c_p, c_p],
    "lue_amd_synthetic_dem_f64": [c_u64, c_i64, c_i64, c_p, c_p],
    "lue_amd_synthetic_uniform_f32": [c_u64, c_u64, c_d, c_d, c_i64, c_p, c_p],
    "lue_amd_synthetic_uniform_f64": [c_u64, c_u64, c_d, c_d, c_i64, c_p, c_p],
}
for _op in ("threshold", "fraction", "capacity", "trigger"):
    SIGNATURES[f"lue_amd_plan_accu_{_op}_f32"] = [c_p, c_p, c_f, c_p, c_f, c_p, c_p, ctypes.POINTER(c_p)]
    SIGNATURES[f"lue_amd_plan_accu_{_op}_f64"] = [c_p, c_p, c_d, c_p, c_d, c_p, c_p, ctypes.POINTER(c_p)]
SIGNATURES["lue_amd_downstream_distance_f32"] = [c_p, c_i64, c_i64, c_f, c_int, c_p, c_p]
SIGNATURES["lue_amd_downstream_distance_f64"] = [c_p, c_i64, c_i64, c_d, c_int, c_p, c_p]
for _op in ("upstream", "downstream"):
    for _sfx in ("u8", "u32", "u64", "i8", "i32", "i64", "f32", "f64"):
        SIGNATURES[f"lue_amd_{_op}_{_sfx}"] = [c_p, c_p, c_i64, c_i64, c_p, c_p]

_RESTYPES = {
    "lue_amd_status_string": ctypes.c_char_p,
    "lue_amd_last_error": ctypes.c_char_p,
    "lue_amd_version": ctypes.c_char_p,
}


STATUS_INVALID_ARGUMENT = 1


class RoutingError(RuntimeError):
    """A C ABI call returned an error status."""

    def __init__(self, status, message):
        super().__init__(message)
        self.status = status


class InvalidNetworkError(RoutingError):
    pass


def load():
    """Load the in-tree library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'` "
            "or `python -m lue_amd.build`")
    lib = ctypes.CDLL(LIB_PATH)
    for name, argtypes in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = argtypes
        fn.restype = _RESTYPES.get(name, c_int)
    _lib = lib
    return lib


SCHEDULES = {"chains": 0, "rounds": 1}


# mirrors of the per-call settings, read by the Python front end
_schedule = "chains"
_network_cache_entries = 4


def set_sweep_options(schedule="chains"):
    """Schedule of the per-call operations after the first round of tiles: "chains" or "rounds"."""
    global _schedule
    check(load().lue_amd_set_sweep_options(SCHEDULES[schedule]))
    _schedule = schedule


def sweep_schedule():
    return _schedule


def set_network_cache(entries=4):
    """Schedules the per-call operations keep (0: record in every call). The C ABI keeps up to `entries`
    schedules found by an ldd fingerprint; the Python front end keeps one per ldd array (0: none)."""
    global _network_cache_entries
    check(load().lue_amd_set_network_cache(entries))
    _network_cache_entries = entries


def network_cache_entries():
    return _network_cache_entries


def trim_scratch():
    """Free the cached per-call device workspaces of the current device that are idle."""
    check(load().lue_amd_trim_scratch())


def profile_enable(enable=True):
    check(load().lue_amd_profile_enable(1 if enable else 0))


def profile_read():
    """(first-round ms, later-rounds ms, calls, rounds, cells) summed since profile_enable()."""
    first, rest = c_d(), c_d()
    calls, rounds, cells = ctypes.c_longlong(), ctypes.c_longlong(), ctypes.c_longlong()
    check(load().lue_amd_profile_read(ctypes.byref(first), ctypes.byref(rest), ctypes.byref(calls),
                                      ctypes.byref(rounds), ctypes.byref(cells)))
    return first.value, rest.value, calls.value, rounds.value, cells.value


OPS = {"accu": 0, "accu_threshold": 1, "accu_fraction": 2, "kinematic_wave": 3, "accu_capacity": 4,
       "accu_trigger": 5}


def profile_read_op(op, f64=False):
    """Per-operation profile since profile_enable(): dict of summed first-round and later-rounds ms, the
    shortest and longest first round, and the number of calls."""
    vals = [c_d() for _ in range(4)]
    calls = ctypes.c_longlong()
    check(load().lue_amd_profile_read_op(OPS[op], 1 if f64 else 0, *[ctypes.byref(v) for v in vals],
                                         ctypes.byref(calls)))
    first, rest, lo, hi = (v.value for v in vals)
    return dict(first_round_ms=first, after_first_round_ms=rest, first_round_min_ms=lo, first_round_max_ms=hi,
                calls=calls.value)


def check(status):
    if status != 0:
        lib = load()
        message = f"{lib.lue_amd_status_string(status).decode()}: {lib.lue_amd_last_error().decode()}"
        if status == 5:
            raise InvalidNetworkError(status, message)
        if status == 2:
            raise ValueError(message)
        raise RoutingError(status, message)
