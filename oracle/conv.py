"""Restated convolution / time-integration semantics of the reference (FP64).

Test infrastructure only (see ``oracle/__init__.py``).

Reference semantics (``pyvoxeldosimetry/core/kernel_convolution.py:71-74``)::

    dose_rate = np.fft.ifftn(np.fft.fftn(activity_map) *
                             np.fft.fftn(self.kernel, activity_map.shape)).real

i.e. a *circular* convolution of size ``activity.shape`` with the kernel
zero-padded at the high end (or cropped to ``[0:N]``) and therefore anchored
at its array origin, not at its centre.

Pinned by (tests/test_oracle.py):
  * ``ref_conv_threaded`` (scipy.fft, threaded) == ``ref_conv`` to FP64 rounding;
  * ``ref_conv`` == ``brute_circular`` (pure Python loops, tiny grids) ==
    ``oracle.brute`` (independent plain C) on seeded inputs, including kernels
    larger than the grid (crop) and even/odd/anisotropic shapes;
  * closed forms: a delta at j returns the padded/cropped kernel rolled by j;
    a constant input c returns c * sum(K') everywhere;
  * ``ref_conv(A, K) == roll(circular_same(A, K), K//2)`` for K <= N;
  * ``linear_same`` == ``scipy.signal.convolve(A, K, 'same', method='direct')``;
  * trapezoid helpers == ``numpy.trapezoid`` on the stacked series.
"""
from __future__ import annotations

import numpy as np

__all__ = [
    "ref_conv",
    "ref_conv_threaded",
    "brute_circular",
    "brute_linear_same",
    "circular_same",
    "linear_same",
    "pad_or_crop",
    "ref_absorbed_dose",
    "trapezoid_weights_seconds",
    "activity_trapezoid",
    "density_factor",
    "epilogue",
    "ref_dose",
    "error_metrics",
]


def pad_or_crop(kernel: np.ndarray, shape) -> np.ndarray:
    """Zero-pad at the high end or crop ``kernel`` to ``shape``.

    This is what ``np.fft.fftn(kernel, s=shape)`` does to its input before the
    transform (``kernel_convolution.py:73``).
    """
    out = np.zeros(tuple(shape), dtype=np.float64)
    sl = tuple(slice(0, min(a, b)) for a, b in zip(kernel.shape, shape))
    out[sl] = kernel[sl]
    return out


def ref_conv(activity: np.ndarray, kernel: np.ndarray) -> np.ndarray:
    """Verbatim restatement of ``kernel_convolution.py:71-74`` (complex128)."""
    a = np.asarray(activity)
    # axes spelled out only to silence NumPy 2's deprecation; identical for 3-D input
    return np.fft.ifftn(np.fft.fftn(a) * np.fft.fftn(kernel, a.shape, axes=(0, 1, 2))).real


def ref_conv_threaded(activity: np.ndarray, kernel: np.ndarray, workers: int = -1) -> np.ndarray:
    """``ref_conv`` with scipy.fft's threaded FP64 transforms, for volumes where the
    single-threaded expression takes too long in a test.  Same operator (circular,
    kernel zero-padded at the high end or cropped); pinned against ``ref_conv``."""
    import scipy.fft as sfft
    a = np.asarray(activity, dtype=np.float64)
    kp = pad_or_crop(np.asarray(kernel, dtype=np.float64), a.shape)
    spec = sfft.rfftn(a, workers=workers) * sfft.rfftn(kp, workers=workers)
    return sfft.irfftn(spec, s=a.shape, workers=workers)


def brute_circular(activity: np.ndarray, kernel: np.ndarray) -> np.ndarray:
    """Direct circular convolution ``D[i] = sum_j A[j] K'[(i-j) mod N]``.

    Pure-Python over kernel taps (vectorised over the grid with ``np.roll``);
    only for small grids.  ``K' = pad_or_crop(K, N)``.
    """
    a = np.asarray(activity, dtype=np.float64)
    kp = pad_or_crop(np.asarray(kernel, dtype=np.float64), a.shape)
    out = np.zeros_like(a)
    # np.nonzero keeps NaN/inf taps (they compare != 0), so they propagate
    for m in zip(*np.nonzero(kp)):
        out += kp[m] * np.roll(a, shift=m, axis=(0, 1, 2))
    return out


def brute_linear_same(activity: np.ndarray, kernel: np.ndarray) -> np.ndarray:
    """Direct centred linear convolution, zero outside the grid.

    ``D[o] = sum_m K[m] * A[o - m + K//2]`` (``A`` zero outside ``[0, N)``).
    """
    a = np.asarray(activity, dtype=np.float64)
    k = np.asarray(kernel, dtype=np.float64)
    n = a.shape
    c = [s // 2 for s in k.shape]
    out = np.zeros_like(a)
    for m in np.ndindex(*k.shape):
        w = k[m]
        if w == 0.0:
            continue
        # o - m + c in [0, N)  ->  o in [m - c, N + m - c)
        src = []
        dst = []
        for ax in range(3):
            sh = c[ax] - m[ax]          # input index = o + sh
            o_lo = max(0, -sh)
            o_hi = min(n[ax], n[ax] - sh)
            if o_lo >= o_hi:
                break
            dst.append(slice(o_lo, o_hi))
            src.append(slice(o_lo + sh, o_hi + sh))
        else:
            out[tuple(dst)] += w * a[tuple(src)]
    return out


def circular_same(activity: np.ndarray, kernel: np.ndarray) -> np.ndarray:
    """Circular convolution with the kernel centred at ``K//2`` (K <= N)."""
    a = np.asarray(activity, dtype=np.float64)
    k = np.asarray(kernel, dtype=np.float64)
    kc = np.zeros(a.shape)
    c = [s // 2 for s in k.shape]
    for m in np.ndindex(*k.shape):
        kc[tuple((m[i] - c[i]) % a.shape[i] for i in range(3))] += k[m]
    return np.fft.ifftn(np.fft.fftn(a) * np.fft.fftn(kc)).real


def linear_same(activity: np.ndarray, kernel: np.ndarray) -> np.ndarray:
    """Centred ('same') linear convolution in FP64 through a padded FFT.

    The physically meaningful mode (no wrap-around); equal to
    ``scipy.signal.fftconvolve(A, K, mode='same')``.
    """
    a = np.asarray(activity, dtype=np.float64)
    k = np.asarray(kernel, dtype=np.float64)
    full = [a.shape[i] + k.shape[i] - 1 for i in range(3)]
    spec = np.fft.rfftn(a, full, axes=(0, 1, 2)) * np.fft.rfftn(k, full, axes=(0, 1, 2))
    lin = np.fft.irfftn(spec, full, axes=(0, 1, 2))
    c = [s // 2 for s in k.shape]
    return lin[c[0]:c[0] + a.shape[0], c[1]:c[1] + a.shape[1], c[2]:c[2] + a.shape[2]]


def ref_absorbed_dose(activity_maps, time_points, kernel, conv=ref_conv) -> np.ndarray:
    """Restatement of ``kernel_convolution.py:78-106``.

    One dose-rate map per timepoint, then
    ``sum_i (D_i + D_{i+1})/2 * (t_{i+1}-t_i)*3600`` (hours -> seconds, no tail).
    Accumulated in FP64 (the reference's ``zeros_like`` follows the input dtype,
    ``:94``; the oracle keeps FP64, the precision every result is checked against).
    """
    rates = [conv(a, kernel) for a in activity_maps]
    dose = np.zeros(np.shape(activity_maps[0]), dtype=np.float64)
    for i in range(len(time_points) - 1):
        dt = (time_points[i + 1] - time_points[i]) * 3600.0
        dose += (rates[i] + rates[i + 1]) / 2.0 * dt
    return dose


def trapezoid_weights_seconds(time_points_h) -> np.ndarray:
    """Per-timepoint weights ``w`` with ``sum_i w_i x_i`` == the trapezoid of
    ``kernel_convolution.py:101-104`` (dt in hours * 3600)."""
    t = np.asarray(time_points_h, dtype=np.float64)
    w = np.zeros(len(t))
    for i in range(len(t) - 1):
        dt = (t[i + 1] - t[i]) * 3600.0
        w[i] += dt / 2.0
        w[i + 1] += dt / 2.0
    return w


def activity_trapezoid(activity_maps, time_points) -> np.ndarray:
    """Restatement of ``core/activity_sampler.py:69-79`` (time in input units)."""
    res = np.zeros(np.shape(activity_maps[0]), dtype=np.float64)
    for i in range(len(time_points) - 1):
        dt = time_points[i + 1] - time_points[i]
        res += (np.asarray(activity_maps[i], np.float64)
                + np.asarray(activity_maps[i + 1], np.float64)) / 2.0 * dt
    return res


def density_factor(density: np.ndarray | None, rho_water: float = 1.0):
    """Declared local-deposition mass correction ``rho_w / rho(x)``.

    The reference accepts ``tissue_densities`` but never reads it
    (``core/dose_calculator.py:90``); this build declares the factor
    ``rho_w/rho`` for ``rho > 0`` and 0 where ``rho <= 0`` (no mass).
    """
    if density is None:
        return 1.0
    d = np.asarray(density, dtype=np.float64)
    out = np.zeros_like(d)
    pos = d > 0
    out[pos] = rho_water / d[pos]
    return out


def epilogue(dose: np.ndarray, density=None, scale: float = 1.0) -> np.ndarray:
    """``out = S * (rho_w/rho(x)) * conv`` in FP64."""
    return np.asarray(dose, np.float64) * density_factor(density) * float(scale)


def ref_dose(activity_maps, kernel, *, time_points=None, weights=None, density=None,
             scale=1.0, mode="reference") -> np.ndarray:
    """Oracle for one fused product call.

    ``mode='reference'`` uses ``ref_conv`` (circular, origin-anchored kernel);
    ``mode='linear'`` uses ``linear_same``.  With several maps the result is the
    trapezoid of per-timepoint dose rates (``ref_absorbed_dose``) or, when
    explicit ``weights`` are given, ``sum_t w_t conv(A_t)``.
    """
    conv = ref_conv if mode == "reference" else linear_same
    maps = activity_maps if isinstance(activity_maps, (list, tuple)) else [activity_maps]
    if weights is not None:
        d = sum(float(w) * conv(a, kernel) for w, a in zip(weights, maps))
    elif len(maps) == 1:
        d = conv(maps[0], kernel)
    else:
        d = ref_absorbed_dose(maps, time_points, kernel, conv=conv)
    return epilogue(d, density, scale)


def error_metrics(test: np.ndarray, ref: np.ndarray) -> dict:
    """``max_rel = max|D-D64|/max|D64|``, ``rms_rel = rms(D-D64)/max|D64|``."""
    t = np.asarray(test, dtype=np.float64)
    r = np.asarray(ref, dtype=np.float64)
    den = float(np.max(np.abs(r))) if r.size else 0.0
    if den == 0.0:
        den = 1.0
    diff = t - r
    return {
        "max_rel": float(np.max(np.abs(diff)) / den) if r.size else 0.0,
        "rms_rel": float(np.sqrt(np.mean(diff * diff)) / den) if r.size else 0.0,
        "max_abs_ref": den,
    }
