"""Grid-to-grid resampling on the GPU (``include/pvd_resample.h``).

``resample`` mirrors ``ImageRegistration.resample_to_reference``
(core/image_registration.py:128-173): output size round(n * in / out) per axis,
origin 0, 0 outside the input, 'linear', 'nearest' or 'cubic' (SimpleITK's
sitkBSpline: the interpolating cubic B-spline with mirror boundaries, its
coefficients from ``pvd_bspline_prefilter``).  Unlike the reference
(which hands a NumPy (z, y, x) array to SimpleITK with an (x, y, z) spacing),
spacing[a] here always belongs to array axis a.  ``resample_activity`` keeps
the total activity of a per-voxel (Bq) map by scaling with the voxel-volume
ratio, which the dose path needs for ``_resample_activity``
(core/kernel_convolution.py:108-115, a stub in the reference).
"""
from __future__ import annotations

import ctypes
from typing import Optional, Sequence

import numpy as np

from . import _native as nat
from .arrays import from_device, to_device_f32

INTERP = {"nearest": 0, "linear": 1, "cubic": 2}


def output_shape_for(shape, input_spacing, output_spacing):
    """round(n * in / out) per axis (image_registration.py:150-153), at least 1."""
    return tuple(max(1, int(round(n * i / o))) for n, i, o in zip(shape, input_spacing, output_spacing))


def resample_device(image, input_spacing: Sequence[float], output_spacing: Sequence[float],
                    interpolation: str = "linear", output_shape: Optional[Sequence[int]] = None, scale: float = 1.0,
                    out=None):
    """GPU tensor in, GPU tensor out (float32)."""
    import torch
    nat.require_device()
    lib = nat.load()
    if interpolation not in INTERP:
        raise ValueError(f"Unsupported interpolation method: {interpolation}")
    if len(input_spacing) != 3 or len(output_spacing) != 3:
        raise ValueError("spacings must have 3 entries")
    x, _ = to_device_f32(image)
    if x.dim() != 3:
        raise ValueError(f"image must be 3-D, got shape {tuple(x.shape)}")
    shape = tuple(output_shape) if output_shape is not None else output_shape_for(x.shape, input_spacing,
                                                                                 output_spacing)
    if out is None:
        out = torch.empty(shape, dtype=torch.float32, device=x.device)
    elif tuple(out.shape) != shape or out.dtype != torch.float32 or not out.is_contiguous():
        raise ValueError("out must be a contiguous float32 tensor of the output shape")
    si = (ctypes.c_double * 3)(*[float(v) for v in input_spacing])
    so = (ctypes.c_double * 3)(*[float(v) for v in output_spacing])
    stream = ctypes.c_void_p(nat.current_stream_handle(x.device))
    src = x
    if interpolation == "cubic":
        # interpolating B-spline coefficients (mirror boundaries) of the input
        src = torch.empty_like(x)
        tmp = torch.empty_like(x)
        with torch.cuda.device(x.device):
            nat.check(lib.pvd_bspline_prefilter(ctypes.c_void_p(x.data_ptr()), *x.shape, ctypes.c_void_p(src.data_ptr()),
                                                ctypes.c_void_p(tmp.data_ptr()), stream), "pvd_bspline_prefilter")
    with torch.cuda.device(x.device):
        nat.check(lib.pvd_resample(ctypes.c_void_p(src.data_ptr()), *x.shape, si, ctypes.c_void_p(out.data_ptr()),
                                   *shape, so, INTERP[interpolation], float(scale), stream), "pvd_resample")
    return out


def resample(image, input_spacing, output_spacing, interpolation: str = "linear", output_shape=None,
             scale: float = 1.0):
    """NumPy in -> NumPy (float32) out; GPU tensor in -> GPU tensor out."""
    _, kind = to_device_f32(image) if not isinstance(image, np.ndarray) else (None, "numpy")
    out = resample_device(image, input_spacing, output_spacing, interpolation, output_shape, scale)
    return from_device(out, kind)


def resample_activity(activity, input_spacing, output_spacing, output_shape=None, interpolation: str = "linear"):
    """Per-voxel activity (Bq/voxel) onto another grid: the concentration is
    interpolated and multiplied by the output voxel volume, so a smooth map keeps
    its total activity."""
    ratio = float(np.prod(output_spacing)) / float(np.prod(input_spacing))
    return resample(activity, input_spacing, output_spacing, interpolation, output_shape, scale=ratio)


def resample_dose_map(dose_map, input_spacing, output_spacing):
    """``resample_dose_map`` (core/utils.py:193-231): order-3 spline resampling onto
    ``output_spacing`` with output shape ceil(n * in / out) per axis; returns
    ``(dose, affine)``, affine = diag(output_spacing + [1]).  Both grids keep the
    corner origin the reference's affines state (``diag(spacing)``); the spline is the
    interpolating cubic B-spline with mirror boundaries, 0 outside [-0.5, n - 0.5)."""
    shape = tuple(max(1, int(np.ceil(n * i / o))) for n, i, o in zip(np.shape(dose_map), input_spacing,
                                                                         output_spacing))
    out = resample(dose_map, input_spacing, output_spacing, "cubic", output_shape=shape)
    affine = np.diag([float(s) for s in output_spacing] + [1.0])
    return out, affine
