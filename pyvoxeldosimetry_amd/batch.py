"""Batched dose convolution of many volumes of uneven sizes.

Each volume keeps its own plan (workspace, kernel spectrum, FFT lengths), so
volumes may differ in grid size, kernel and mode.  Per FFT pass, the volumes
that run the same kernel instance with the same block share one launch whose
grid concatenates their tiles (``pvd_batch_*`` in include/pvd.h); a batched
volume therefore gives the bits of its single-volume plan.  ``run(graph=True)``
replays the launches from a hipGraph captured on first use.
"""
from __future__ import annotations

import ctypes
from typing import List, Optional, Sequence

from . import _native as nat
from .plan import ConvPlan, _as_device_f32


class DoseBatch:
    def __init__(self, shapes: Sequence[Sequence[int]], kernels: Sequence, mode: str = "reference",
                 fft_n: Optional[Sequence] = None, device=None, tile: Optional[Sequence[int]] = None,
                 smooth_lengths: bool = False):
        """``fft_n``: per-volume FFT lengths (None entries: the plan's choice); ``tile``:
        lines per FFT tile of the (x, y, z) passes of every volume (None: the plan default)."""
        import torch
        nat.require_device()
        if len(shapes) != len(kernels) or not shapes:
            raise ValueError("need one kernel per volume and at least one volume")
        self.lib = nat.load()
        self.plans: List[ConvPlan] = []
        for i, (shp, k) in enumerate(zip(shapes, kernels)):
            fn = fft_n[i] if fft_n else None
            plan = ConvPlan(shp, tuple(k.shape), mode=mode, algo="fft", fft_n=fn, device=device, tile=tile,
                            smooth_lengths=smooth_lengths)
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
