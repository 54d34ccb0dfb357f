"""Python handle on a ``pvd_plan`` (include/pvd.h): one grid/kernel shape.

Device memory and streams come from PyTorch-ROCm; the arithmetic is the HIP
code in ``csrc/``.  Inputs are torch tensors on the plan's device (float32,
contiguous, C order ``(nx, ny, nz)``); outputs are float32 tensors.
"""
from __future__ import annotations

import ctypes
from typing import Optional, Sequence

from . import _native as nat


def _as_device_f32(t, shape, device, what):
    import torch
    if not isinstance(t, torch.Tensor):
        raise TypeError(f"{what} must be a torch.Tensor on {device}")
    if t.device != device:
        raise ValueError(f"{what} is on {t.device}, plan is on {device}")
    if t.dtype != torch.float32:
        raise ValueError(f"{what} must be float32, got {t.dtype}")
    if tuple(t.shape) != tuple(shape):
        raise ValueError(f"{what} has shape {tuple(t.shape)}, plan expects {tuple(shape)}")
    if not t.is_contiguous():
        raise ValueError(f"{what} must be contiguous")
    return t


def _params(shape, kernel_shape, mode, algo, fft_n, flags, tile, max_radix):
    if mode not in nat.MODES:
        raise ValueError(f"mode must be one of {list(nat.MODES)}, got {mode!r}")
    if algo not in nat.ALGOS:
        raise ValueError(f"algo must be one of {list(nat.ALGOS)}, got {algo!r}")
    if len(shape) != 3 or len(kernel_shape) != 3:
        raise ValueError("shape and kernel_shape must have 3 entries")
    prm = nat.PlanParams()
    prm.n[:] = tuple(int(s) for s in shape)
    prm.k[:] = tuple(int(s) for s in kernel_shape)
    prm.mode = nat.MODES[mode]
    prm.algo = nat.ALGOS[algo]
    prm.fft_n[:] = tuple(int(x) for x in fft_n) if fft_n else (0, 0, 0)
    prm.flags = int(flags)
    prm.tile[:] = tuple(int(t) for t in tile) if tile else (0, 0, 0)
    prm.max_radix = int(max_radix)
    return prm


def _info_dict(info) -> dict:
    fft = info.algo == nat.PVD_ALGO_FFT
    return {"algo": {nat.PVD_ALGO_FFT: "fft", nat.PVD_ALGO_DIRECT: "direct"}[info.algo],
            "fft_n": tuple(info.fft_n), "kernel_used": tuple(info.kernel_used), "n_passes": info.n_passes,
            "workspace_bytes": int(info.workspace_bytes), "tile": tuple(info.tile),
            "max_radix": int(info.max_radix),
            "fft_path": tuple(nat.PATHS[p] for p in info.fft_path) if fft else None,
            "block": tuple(info.block), "grid": tuple(info.grid)}


def describe_plan(shape: Sequence[int], kernel_shape: Sequence[int], mode: str = "reference", algo: str = "auto",
                  fft_n: Optional[Sequence[int]] = None, tile: Optional[Sequence[int]] = None, max_radix: int = 0,
                  stockham: bool = False, smooth_lengths: bool = False) -> dict:
    """The plan ``ConvPlan`` would build for these arguments on a gfx950 device (FFT
    lengths, algorithm, transform path per axis, tiles, launch shapes), computed on the
    host by ``pvd_plan_describe`` without touching a device."""
    lib = nat.load()
    flags = (nat.PVD_FLAG_STOCKHAM if stockham else 0) | (nat.PVD_FLAG_SMOOTH_LENGTHS if smooth_lengths else 0)
    prm = _params(shape, kernel_shape, mode, algo, fft_n, flags, tile, max_radix)
    info = nat.PlanInfo()
    nat.check(lib.pvd_plan_describe(ctypes.byref(prm), ctypes.byref(info)), "pvd_plan_describe")
    return _info_dict(info)


class ConvPlan:
    """Plan for dose = scale * (rho_w/rho) * conv(activity, kernel)."""

    def __init__(self, shape: Sequence[int], kernel_shape: Sequence[int], mode: str = "reference",
                 algo: str = "auto", fft_n: Optional[Sequence[int]] = None, profile: bool = False,
                 device=None, general_spectrum: bool = False, tile: Optional[Sequence[int]] = None,
                 max_radix: int = 0, stockham: bool = False, smooth_lengths: bool = False):
        """``tile``: lines per FFT tile of the (x, y, z) passes (powers of two, 0 = 16, but 8
        on x for the two-step lengths 288..512);
        ``max_radix``: largest power-of-two Stockham radix (16 or 8, 0 = 16);
        ``stockham``: LDS Stockham stages on every axis instead of the register-resident
        transforms (256-point y/x, two-step y/x/z);
        ``smooth_lengths``: automatic FFT lengths from the 7-smooth rule only (by default a
        length is padded further when that reaches a register-resident transform)."""
        import torch
        nat.require_device()
        self.lib = nat.load()
        flags = ((nat.PVD_FLAG_PROFILE if profile else 0) | (nat.PVD_FLAG_GENERAL_SPECTRUM if general_spectrum else 0)
                 | (nat.PVD_FLAG_STOCKHAM if stockham else 0) | (nat.PVD_FLAG_SMOOTH_LENGTHS if smooth_lengths else 0))
        prm = _params(shape, kernel_shape, mode, algo, fft_n, flags, tile, max_radix)
        self.device = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
        self.shape = tuple(int(s) for s in shape)
        self.kernel_shape = tuple(int(s) for s in kernel_shape)
        self.mode = mode
        self.stockham = bool(stockham)
        h = ctypes.c_void_p()
        with torch.cuda.device(self.device):
            nat.check(self.lib.pvd_plan_create(ctypes.byref(h), ctypes.byref(prm)), "pvd_plan_create")
        self._h = h
        info = nat.PlanInfo()
        nat.check(self.lib.pvd_plan_info_get(self._h, ctypes.byref(info)), "pvd_plan_info_get")
        self.algo = {nat.PVD_ALGO_FFT: "fft", nat.PVD_ALGO_DIRECT: "direct"}[info.algo]
        self.fft_n = tuple(info.fft_n)
        self.kernel_used = tuple(info.kernel_used)
        self.n_passes = info.n_passes
        self.workspace_bytes = int(info.workspace_bytes)
        self.tile = tuple(info.tile)
        self.max_radix = int(info.max_radix)
        self.fft_path = _info_dict(info)["fft_path"]
        self.symmetric_spectrum = False
        self._kernel_dev = None

    # ------------------------------------------------------------------
    def info(self) -> dict:
        return {"algo": self.algo, "mode": self.mode, "fft_n": self.fft_n, "kernel_used": self.kernel_used,
                "n_passes": self.n_passes, "workspace_bytes": self.workspace_bytes,
                "symmetric_spectrum": self.symmetric_spectrum, "tile": self.tile, "max_radix": self.max_radix,
                "fft_path": self.fft_path}

    def set_kernel(self, kernel) -> None:
        """FP64 dose kernel (torch tensor on the device or anything numpy accepts)."""
        import numpy as np
        import torch
        if isinstance(kernel, torch.Tensor):
            k = kernel.to(device=self.device, dtype=torch.float64).contiguous()
        else:
            k = torch.from_numpy(np.ascontiguousarray(kernel, dtype=np.float64)).to(self.device)
        if tuple(k.shape) != self.kernel_shape:
            raise ValueError(f"kernel shape {tuple(k.shape)} != plan kernel shape {self.kernel_shape}")
        with torch.cuda.device(self.device):
            nat.check(self.lib.pvd_plan_set_kernel(self._h, ctypes.c_void_p(k.data_ptr()),
                                                   ctypes.c_void_p(nat.current_stream_handle(self.device))),
                      "pvd_plan_set_kernel")
        self._kernel_dev = k
        info = nat.PlanInfo()
        nat.check(self.lib.pvd_plan_info_get(self._h, ctypes.byref(info)), "pvd_plan_info_get")
        self.symmetric_spectrum = bool(info.symmetric_spectrum)

    def _prep(self, density, out):
        import torch
        d = None
        if density is not None:
            d = _as_device_f32(density, self.shape, self.device, "density")
        if out is None:
            out = torch.empty(self.shape, dtype=torch.float32, device=self.device)
        else:
            _as_device_f32(out, self.shape, self.device, "out")
        return d, out

    def dose(self, activity, density=None, scale: float = 1.0, out=None):
        """One dose-rate map: ``scale * (rho_w/rho) * conv(activity)``."""
        import torch
        a = _as_device_f32(activity, self.shape, self.device, "activity")
        d, out = self._prep(density, out)
        with torch.cuda.device(self.device):
            nat.check(self.lib.pvd_conv_dose(self._h, ctypes.c_void_p(a.data_ptr()),
                                             ctypes.c_void_p(d.data_ptr() if d is not None else 0),
                                             ctypes.c_void_p(out.data_ptr()), float(scale),
                                             ctypes.c_void_p(nat.current_stream_handle(self.device))),
                      "pvd_conv_dose")
        return out

    def _acts(self, activities):
        if not 1 <= len(activities) <= nat.PVD_MAX_TIMEPOINTS:
            raise ValueError(f"need 1..{nat.PVD_MAX_TIMEPOINTS} activity maps, got {len(activities)}")
        acts = [_as_device_f32(a, self.shape, self.device, f"activity[{i}]") for i, a in enumerate(activities)]
        arr = (ctypes.c_void_p * len(acts))(*[a.data_ptr() for a in acts])
        return acts, arr

    def dose_weighted(self, activities, weights, density=None, scale: float = 1.0, out=None):
        """``scale * (rho_w/rho) * conv(sum_t w_t A_t)``."""
        import torch
        acts, arr = self._acts(activities)
        if len(weights) != len(acts):
            raise ValueError("one weight per activity map")
        w = (ctypes.c_double * len(acts))(*[float(x) for x in weights])
        d, out = self._prep(density, out)
        with torch.cuda.device(self.device):
            nat.check(self.lib.pvd_conv_dose_weighted(self._h, arr, w, len(acts),
                                                      ctypes.c_void_p(d.data_ptr() if d is not None else 0),
                                                      ctypes.c_void_p(out.data_ptr()), float(scale),
                                                      ctypes.c_void_p(nat.current_stream_handle(self.device))),
                      "pvd_conv_dose_weighted")
        return out

    def dose_timeseries(self, activities, time_points_h, density=None, scale: float = 1.0, out=None):
        """Trapezoid time-integrated dose (hours, x3600) in one convolution."""
        import torch
        acts, arr = self._acts(activities)
        if len(time_points_h) != len(acts):
            raise ValueError("Number of activity maps must match number of time points")
        t = (ctypes.c_double * len(acts))(*[float(x) for x in time_points_h])
        d, out = self._prep(density, out)
        with torch.cuda.device(self.device):
            nat.check(self.lib.pvd_conv_dose_timeseries(self._h, arr, t, len(acts),
                                                        ctypes.c_void_p(d.data_ptr() if d is not None else 0),
                                                        ctypes.c_void_p(out.data_ptr()), float(scale),
                                                        ctypes.c_void_p(nat.current_stream_handle(self.device))),
                      "pvd_conv_dose_timeseries")
        return out

    def timings(self) -> dict:
        """Per-pass GPU ms of the last call (plan created with profile=True)."""
        ms = (ctypes.c_float * 8)()
        n = nat.check(self.lib.pvd_plan_timings(self._h, ms, 8), "pvd_plan_timings")
        return {self.lib.pvd_pass_name(self._h, i).decode(): float(ms[i]) for i in range(n)}

    def close(self) -> None:
        if getattr(self, "_h", None):
            self.lib.pvd_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
