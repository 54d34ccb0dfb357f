"""Activity time integration on the GPU.

Mirrors ``ActivitySampler`` (``pyvoxeldosimetry/core/activity_sampler.py:8-79``):
same constructor checks (``half_life > 0``, units in hours/minutes/seconds),
``sample_timepoints`` and ``integrate_activity`` (trapezoid with the time step
in the caller's units, no conversion).  The trapezoid runs as one
``pvd_weighted_sum`` launch.  ``integrate_dose_rates`` is the method the
reference's ``DoseCalculator`` calls in its dose-rate mode
(``core/dose_calculator.py:138``) but never defines; it is provided here.
"""
from __future__ import annotations

import ctypes
import math
from typing import List, Optional

import numpy as np

from .. import _native as nat
from ..arrays import to_device_f32, from_device

_UNIT_SECONDS = {"hours": 3600.0, "minutes": 60.0, "seconds": 1.0}


def trapezoid_weights(time_points) -> List[float]:
    """Weights w with sum_i w_i x_i == sum_i (x_i + x_{i+1})/2 * (t_{i+1} - t_i)."""
    t = [float(x) for x in time_points]
    w = [0.0] * len(t)
    for i in range(len(t) - 1):
        dt = t[i + 1] - t[i]
        w[i] += dt / 2.0
        w[i + 1] += dt / 2.0
    return w


def weighted_sum_device(maps, weights, out=None):
    """out = sum_t w_t maps[t] on the GPU (all maps float32 device tensors of one shape)."""
    import torch
    nat.require_device()
    lib = nat.load()
    if not 1 <= len(maps) <= nat.PVD_MAX_TIMEPOINTS:
        raise ValueError(f"need 1..{nat.PVD_MAX_TIMEPOINTS} maps, got {len(maps)}")
    shape = tuple(maps[0].shape)
    dev = maps[0].device
    for m in maps:
        if tuple(m.shape) != shape or m.dtype != torch.float32 or m.device != dev or not m.is_contiguous():
            raise ValueError("maps must be contiguous float32 tensors of one shape on one device")
    if out is None:
        out = torch.empty(shape, dtype=torch.float32, device=dev)
    arr = (ctypes.c_void_p * len(maps))(*[m.data_ptr() for m in maps])
    w = (ctypes.c_double * len(maps))(*[float(x) for x in weights])
    with torch.cuda.device(dev):
        nat.check(lib.pvd_weighted_sum(arr, w, len(maps), ctypes.c_size_t(out.numel()),
                                       ctypes.c_void_p(out.data_ptr()),
                                       ctypes.c_void_p(nat.current_stream_handle(dev))), "pvd_weighted_sum")
    return out


class ActivitySampler:
    def __init__(self, half_life: float, units: str = "hours"):
        self.half_life = half_life
        self.units = units
        self._validate_inputs()

    def _validate_inputs(self):
        if self.half_life <= 0:
            raise ValueError("Half-life must be positive")
        if self.units not in _UNIT_SECONDS:
            raise ValueError("Invalid time units")

    @property
    def decay_constant(self) -> float:
        """lambda = ln 2 / T_half, per unit of ``units``."""
        return math.log(2.0) / self.half_life

    def sample_timepoints(self, start_time: float, end_time: float, n_points: int = 10) -> np.ndarray:
        return np.linspace(start_time, end_time, n_points)

    def integrate_activity(self, activity_maps, time_points, method: str = "trapezoid"):
        """Trapezoid over the maps, time step in ``units`` (``activity_sampler.py:48-79``)."""
        if method != "trapezoid":
            raise ValueError(f"Unknown integration method: {method}")
        maps, kind = zip(*[to_device_f32(a) for a in activity_maps])
        out = weighted_sum_device(list(maps), trapezoid_weights(time_points))
        return from_device(out, kind[0])

    def integrate_dose_rates(self, dose_rates, time_points, integration_limit: Optional[float] = None):
        """Trapezoid over dose-rate maps with dt converted to seconds; beyond the
        last timepoint, if ``integration_limit`` (same units) is later, add the
        physical-decay tail ``D_last (1 - e^{-lambda (L - t_last)}) / lambda``
        (the mono-exponential tail of ``time_integration/curve_fitting.py:74-84``)."""
        to_s = _UNIT_SECONDS[self.units]
        w = [x * to_s for x in trapezoid_weights(time_points)]
        if integration_limit is not None and integration_limit > time_points[-1]:
            lam = self.decay_constant
            span = float(integration_limit) - float(time_points[-1])
            w[-1] += (1.0 - math.exp(-lam * span)) / lam * to_s
        maps, kind = zip(*[to_device_f32(a) for a in dose_rates])
        out = weighted_sum_device(list(maps), w)
        return from_device(out, kind[0])
