"""Batched dose convolution of many volumes of uneven sizes.

All volumes go through one launch per FFT pass (``pvd_batch_*`` in
include/pvd.h); ``run(graph=True)`` replays the five launches from a hipGraph
captured on first use.  Each volume keeps its own plan (workspace and kernel
spectrum), so volumes may differ in grid size, kernel and mode.
"""
from __future__ import annotations

import ctypes
from typing import List, Optional, Sequence

from . import _native as nat
from .plan import ConvPlan, _as_device_f32


# Default lines per FFT tile (x, y, z) of batched volumes: on 48 phantoms of 96^3 to
# 384x384x320 (T = 4) 8-line x/y tiles ran the batch in 12.37 ms against 13.22 ms with
# the single-volume default of 16 (16/8/16: 12.99, 8/8/8: 12.45, 8/4/16: 13.64 ms).
# Axes of FFT length 256 keep 16-line tiles: their register-resident path needs them,
# and with it a batched volume gives the same bits as its single-volume plan.
BATCH_TILE = (8, 8, 16)


class DoseBatch:
    def __init__(self, shapes: Sequence[Sequence[int]], kernels: Sequence, mode: str = "reference",
                 fft_n: Optional[Sequence] = None, device=None, tile: Optional[Sequence[int]] = None):
        """``tile``: lines per FFT tile of the (x, y, z) passes of every volume (BATCH_TILE
        when not given; 0 entries = the single-volume default)."""
        import torch
        nat.require_device()
        if len(shapes) != len(kernels) or not shapes:
            raise ValueError("need one kernel per volume and at least one volume")
        self.lib = nat.load()
        self.plans: List[ConvPlan] = []
        for i, (shp, k) in enumerate(zip(shapes, kernels)):
            fn = fft_n[i] if fft_n else None
            if tile is not None:
                plan = ConvPlan(shp, tuple(k.shape), mode=mode, algo="fft", fft_n=fn, device=device, tile=tile)
            else:
                plan = ConvPlan(shp, tuple(k.shape), mode=mode, algo="fft", fft_n=fn, device=device)
                t = tuple(16 if plan.fft_n[a] == 256 else BATCH_TILE[a] for a in range(2)) + (BATCH_TILE[2],)
                if t != (16, 16, 16):                   # the plan's own default
                    plan.close()
                    plan = ConvPlan(shp, tuple(k.shape), mode=mode, algo="fft", fft_n=fn, device=device, tile=t)
            plan.set_kernel(k)
            self.plans.append(plan)
        self.device = self.plans[0].device
        arr = (ctypes.c_void_p * len(self.plans))(*[p._h.value for p in self.plans])
        h = ctypes.c_void_p()
        with torch.cuda.device(self.device):
            nat.check(self.lib.pvd_batch_create(ctypes.byref(h), arr, len(self.plans)), "pvd_batch_create")
        self._h = h
        self._keep = [None] * len(self.plans)

    def __len__(self):
        return len(self.plans)

    def set_volume(self, i: int, activities, weights, density=None, out=None, scale: float = 1.0):
        """Inputs of volume i: dose = scale * (rho_w/rho) * conv(sum_t w_t A_t)."""
        import torch
        if not 0 <= i < len(self.plans):
            raise ValueError(f"volume index {i} outside [0, {len(self.plans)})")
        plan = self.plans[i]
        if len(activities) != len(weights) or not 1 <= len(activities) <= nat.PVD_MAX_TIMEPOINTS:
            raise ValueError("one weight per activity map, 1..16 maps")
        acts = [_as_device_f32(a, plan.shape, plan.device, f"activity[{t}]") for t, a in enumerate(activities)]
        dens = _as_device_f32(density, plan.shape, plan.device, "density") if density is not None else None
        if out is None:
            out = torch.empty(plan.shape, dtype=torch.float32, device=plan.device)
        else:
            _as_device_f32(out, plan.shape, plan.device, "out")
        arr = (ctypes.c_void_p * len(acts))(*[a.data_ptr() for a in acts])
        w = (ctypes.c_double * len(acts))(*[float(x) for x in weights])
        nat.check(self.lib.pvd_batch_set_volume(self._h, i, arr, w, len(acts),
                                                ctypes.c_void_p(dens.data_ptr() if dens is not None else 0),
                                                ctypes.c_void_p(out.data_ptr()), float(scale)),
                  "pvd_batch_set_volume")
        self._keep[i] = (acts, dens, out)      # keep the tensors alive while the batch uses them
        return out

    def set_volume_timeseries(self, i: int, activities, time_points_h, density=None, out=None, scale: float = 1.0):
        """Trapezoid time-integrated dose (hours x 3600), kernel_convolution.py:78-106 semantics."""
        from .core.activity_sampler import trapezoid_weights
        w = [x * 3600.0 for x in trapezoid_weights(time_points_h)]
        return self.set_volume(i, activities, w, density, out, scale)

    def run(self, graph: bool = True):
        """Launch the whole batch on the current stream; returns the output tensors."""
        import torch
        fn = self.lib.pvd_batch_run_graph if graph else self.lib.pvd_batch_run
        with torch.cuda.device(self.device):
            nat.check(fn(self._h, ctypes.c_void_p(nat.current_stream_handle(self.device))), "pvd_batch_run")
        return [k[2] if k else None for k in self._keep]

    def close(self):
        if getattr(self, "_h", None):
            self.lib.pvd_batch_destroy(self._h)
            self._h = None
        for p in getattr(self, "plans", []):
            p.close()
        self.plans = []

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
