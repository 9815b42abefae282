"""Time-dependent targets and scalar read-back (misc/make_function.py,
misc/extract_scalars.py).

The functions returned by ``make_function_*`` are ordinary Python callables
(as in the reference) that also carry a descriptor, so the integrator kernels
can evaluate the same schedule on the device.
"""
import math

import numpy as np


def make_function_constant(value):
    """make_function.py:6-13"""
    value = np.float32(value)

    def function(x):
        return value

    function._gmd_schedule = (0, [float(value), 0.0, 0.0, 0.0])
    return function


def make_function_ramp(value0, x0, value1, x1):
    """make_function.py:16-34: value0 for x < x0, linear to value1 at x1, then value1."""
    value0, x0, value1, x1 = np.float32(value0), np.float32(x0), np.float32(value1), np.float32(x1)
    alpha = (value1 - value0) / (x1 - x0)

    def function(x):
        if x < x0:
            return value0
        if x < x1:
            return value0 + (x - x0) * alpha
        return value1

    function._gmd_schedule = (1, [float(value0), float(x0), float(value1), float(x1)])
    return function


def make_function_sin(period, amplitude, offset):
    """make_function.py:37-47"""
    period, amplitude, offset = np.float32(period), np.float32(amplitude), np.float32(offset)

    def function(x):
        return offset + amplitude * math.sin(2 * math.pi * x / period)

    function._gmd_schedule = (2, [float(period), float(amplitude), float(offset), 0.0])
    return function


def temperature_schedule(temperature):
    """(kind, params) of a temperature given as a number or a make_function_* callable."""
    if callable(temperature):
        sched = getattr(temperature, "_gmd_schedule", None)
        if sched is None:
            raise ValueError("temperature functions must come from make_function_constant/ramp/sin "
                             "to run on the MI355X kernels")
        return sched
    return 0, [float(np.float32(temperature)), 0.0, 0.0, 0.0]


def extract_scalars(data, column_list, first_block=0, D=3):
    """misc/extract_scalars.py:1-51: flattened per-save columns from sim.output."""
    column_indices = {}
    try:
        scalar_names = data.attrs['scalar_names']
    except KeyError:
        scalar_names = data.attrs['scalars_names']
    for index, name in enumerate(scalar_names):
        column_indices[name] = index
    output_list = []
    for column in column_list:
        output_list.append(np.asarray(data['scalars'])[first_block:, :, column_indices[column]].flatten())
    return tuple(output_list)
