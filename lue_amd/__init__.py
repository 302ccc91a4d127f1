"""MI355X-native LUE flow-routing path: accu, accu_threshold, accu_fraction, accu_capacity, accu_trigger,
kinematic_wave, inflow_count.

The compute runs in hand-written gfx950 HIP kernels (liblue_amd.so, C ABI in include/lue_amd/routing.h).
`lue_amd.framework` mirrors lue.framework's routing operations on device-resident PartitionedArrays.
"""
from . import _lib  # noqa: F401

__all__ = ["framework", "synthetic"]
