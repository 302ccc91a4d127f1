"""PCRaster-style names for the routing operations, mirroring lue.pcraster's routing subset.

Reference: source/framework/python/pcraster/__init__.py:457-861 (accu*, catchmenttotal, downstream,
downstreamdist, kinematic, ldd, lddcreate, scalar, upstream). As there, every material argument is converted to
the scalar value scale (float32): a float32 PartitionedArray is used as is, another PartitionedArray is cast
to float32 (its no-data becomes NaN), a number becomes a float32 Scalar; the flow direction is converted to
the ldd value scale (uint8). The state / flux pairs are the two outputs of one lue.framework call.

Operations the reference leaves unimplemented raise NotImplementedError here too (accutraveltime*,
kinwavestate, kinwaveflux, lddcreatedem, ldddist, lddmask, lddrepair). Reading rasters from file names is
not available: arguments are PartitionedArrays or numbers.
"""
from __future__ import annotations

import numpy as np
import torch

from . import framework as lfr


class Configuration:
    """lue.pcraster.configuration: the cell size used by downstreamdist, cellarea and celllength."""

    def __init__(self):
        self.cell_size = 1.0


configuration = Configuration()


def _is_value(argument):
    return isinstance(argument, (int, float, np.integer, np.floating)) and not isinstance(argument, bool)


def _read_if_necessary(argument):
    if isinstance(argument, str):
        raise NotImplementedError("reading rasters from file names is not supported: pass a PartitionedArray")
    return argument


def _integer_no_data(dtype):
    info = torch.iinfo(dtype)
    return info.min if info.min < 0 else info.max


def _cast(array, dtype):
    """lfr.cast under value policies: values converted, the input's no-data becomes the output's."""
    data = array.data
    if data.dtype.is_floating_point:
        no_data = torch.isnan(data)
    else:
        no_data = data == _integer_no_data(data.dtype)
    out = data.to(dtype)
    if dtype.is_floating_point:
        out = out.masked_fill(no_data, float("nan"))
    else:
        out = out.masked_fill(no_data, _integer_no_data(dtype))
    return lfr.PartitionedArray(out, array.partition_shape)


def scalar(expression):
    """Scalar value scale (float32): pcraster/__init__.py:1064-1074."""
    expression = _read_if_necessary(expression)
    if isinstance(expression, lfr.PartitionedArray):
        return expression if expression.dtype == torch.float32 else _cast(expression, torch.float32)
    if _is_value(expression):
        return float(np.float32(expression))
    raise RuntimeError(f"Unsupported argument: {expression}")


def ldd(expression):
    """Ldd value scale (uint8): pcraster/__init__.py:842-852."""
    expression = _read_if_necessary(expression)
    if isinstance(expression, lfr.PartitionedArray):
        return expression if expression.dtype == torch.uint8 else _cast(expression, torch.uint8)
    if _is_value(expression):
        raise TypeError("a non-spatial ldd is not supported by the routing operations")
    raise RuntimeError(f"Unsupported argument: {expression}")


def _spatial(value, template):
    """A float32 raster of `value` shaped and partitioned like `template` (non_spatial_to_spatial)."""
    return lfr.create_partitioned_array(template.shape, template.partition_shape, torch.float32, float(value),
                                        device=template.data.device)


def accuflux(flow_direction, inflow):
    return lfr.accu(ldd(flow_direction), scalar(inflow))


def accucapacity(flow_direction, inflow, capacity):
    return lfr.accu_capacity(ldd(flow_direction), scalar(inflow), scalar(capacity))


def accucapacityflux(flow_direction, inflow, capacity):
    return accucapacity(flow_direction, inflow, capacity)[0]


def accucapacitystate(flow_direction, inflow, capacity):
    return accucapacity(flow_direction, inflow, capacity)[1]


def accufraction(flow_direction, inflow, fraction):
    return lfr.accu_fraction(ldd(flow_direction), scalar(inflow), scalar(fraction))


def accufractionflux(flow_direction, inflow, fraction):
    return accufraction(flow_direction, inflow, fraction)[0]


def accufractionstate(flow_direction, inflow, fraction):
    return accufraction(flow_direction, inflow, fraction)[1]


def accuthreshold(flow_direction, inflow, threshold):
    return lfr.accu_threshold(ldd(flow_direction), scalar(inflow), scalar(threshold))


def accuthresholdflux(flow_direction, inflow, threshold):
    return accuthreshold(flow_direction, inflow, threshold)[0]


def accuthresholdstate(flow_direction, inflow, threshold):
    return accuthreshold(flow_direction, inflow, threshold)[1]


def accutrigger(flow_direction, inflow, trigger):
    return lfr.accu_trigger(ldd(flow_direction), scalar(inflow), scalar(trigger))


def accutriggerflux(flow_direction, inflow, trigger):
    return accutrigger(flow_direction, inflow, trigger)[0]


def accutriggerstate(flow_direction, inflow, trigger):
    return accutrigger(flow_direction, inflow, trigger)[1]


def catchmenttotal(amount, flow_direction):
    """The remainder of accuthreshold with threshold 0 (pcraster/__init__.py:634-637)."""
    return accuthreshold(flow_direction, amount, 0)[1]


def kinematic(flow_direction, current_outflow, inflow, alpha, beta, nr_time_slices, time_step_duration,
              channel_length):
    """lue.framework.kinematic_wave; nr_time_slices is not used (pcraster/__init__.py:804-831)."""
    del nr_time_slices
    flow_direction = ldd(flow_direction)
    current_outflow = scalar(current_outflow)
    inflow = scalar(inflow)
    # current outflow and lateral inflow are rasters in lue.framework.kinematic_wave
    if not isinstance(current_outflow, lfr.PartitionedArray):
        current_outflow = _spatial(current_outflow, flow_direction)
    if not isinstance(inflow, lfr.PartitionedArray):
        inflow = _spatial(inflow, flow_direction)
    time_step_duration = scalar(time_step_duration)
    if isinstance(time_step_duration, lfr.PartitionedArray):
        raise TypeError("time_step_duration must be non-spatial")
    return lfr.kinematic_wave(flow_direction, current_outflow, inflow, scalar(alpha), scalar(beta),
                              time_step_duration, scalar(channel_length))


def upstream(flow_direction, material):
    """pcraster/__init__.py:1245-1252: a non-spatial material becomes a float32 raster first."""
    flow_direction = _read_if_necessary(flow_direction)
    material = _read_if_necessary(material)
    if _is_value(material):
        material = _spatial(np.float32(material), flow_direction)
    return lfr.upstream(flow_direction, material)


def downstream(flow_direction, expression):
    """pcraster/__init__.py:696-699."""
    return lfr.downstream(_read_if_necessary(flow_direction), _read_if_necessary(expression))


def downstreamdist(flow_direction):
    """pcraster/__init__.py:702-705: float32 distances with configuration.cell_size."""
    return lfr.downstream_distance(_read_if_necessary(flow_direction), configuration.cell_size, np.float32)


def lddcreate(elevation, outflowdepth, corevolume, corearea, catchmentprecipitation):
    """D8 flow direction of the elevation; the pit-removal arguments are not used (pcraster/__init__.py:855-858)."""
    del outflowdepth, corevolume, corearea, catchmentprecipitation
    return lfr.d8_flow_direction(_read_if_necessary(elevation))


def cellarea(*args):
    return configuration.cell_size ** 2


def celllength(*args):
    return configuration.cell_size


def _not_implemented(name):
    def operation(*args, **kwargs):
        raise NotImplementedError(name)

    operation.__name__ = name
    return operation


accutraveltimeflux = _not_implemented("accutraveltimeflux")
accutraveltimestate = _not_implemented("accutraveltimestate")
accutraveltimefractionflux = _not_implemented("accutraveltimefractionflux")
accutraveltimefractionstate = _not_implemented("accutraveltimefractionstate")
kinwavestate = _not_implemented("kinwavestate")
kinwaveflux = _not_implemented("kinwaveflux")
lddcreatedem = _not_implemented("lddcreatedem")
ldddist = _not_implemented("ldddist")
lddmask = _not_implemented("lddmask")
lddrepair = _not_implemented("lddrepair")
