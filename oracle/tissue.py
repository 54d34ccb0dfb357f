"""CPU oracle of the HU -> tissue path (test infrastructure only; never on the
product path).

Restates pyvoxeldosimetry/tissue/composition.py:48-94 with the same library
calls (NumPy masks, scipy.ndimage.gaussian_filter with its defaults), and the
build's declared HU -> material -> density map.  FP64 throughout.

Pinned in tests/test_oracle_tissue.py by: the identity without metal; a single
metal voxel in a constant field (closed form c*(1 - w0^3)); a brute-force
(2r+1)^3 sum with explicitly reflected indices on small grids, including
grids smaller than the filter radius; the material table against the
reference's own data files (tests/golden/hu_materials.json).
"""
from __future__ import annotations

import numpy as np
from scipy.ndimage import gaussian_filter

CLASSES = {"air": (-1000, -900), "lung": (-900, -500), "soft_tissue": (-100, 100), "bone": (300, 3000),
           "water": (-10, 10)}


def handle_artifacts(ct, metal_threshold=2000.0, sigma=1.0):
    """composition.py:72-94: metal -> NaN, Gaussian of nan_to_num, metal <- smoothed."""
    corrected = np.array(ct, dtype=np.float64, copy=True)
    metal = corrected > metal_threshold
    corrected[metal] = np.nan
    smoothed = gaussian_filter(np.nan_to_num(corrected), sigma=sigma)
    corrected[metal] = smoothed[metal]
    return corrected


def calculate_composition(ct, handle=True, metal_threshold=2000.0, sigma=1.0):
    """composition.py:48-70: inclusive range masks as float64."""
    img = handle_artifacts(ct, metal_threshold, sigma) if handle else np.asarray(ct, dtype=np.float64)
    return {name: ((img >= lo) & (img <= hi)).astype(float) for name, (lo, hi) in CLASSES.items()}


def gaussian_weights(sigma=1.0, truncate=4.0):
    """scipy's 1-D Gaussian: radius int(truncate*sigma + 0.5), exp(-x^2/(2 sigma^2)) normalised."""
    r = int(truncate * sigma + 0.5)
    x = np.arange(-r, r + 1, dtype=np.float64)
    w = np.exp(-0.5 / sigma ** 2 * x ** 2)
    return w / w.sum()


def reflect_index(i, n):
    """scipy.ndimage 'reflect' (d c b a | a b c d | d c b a), any number of folds."""
    p = 2 * n
    i = i % p
    return i if i < n else p - 1 - i


def brute_fill_at(ct, idx, metal_threshold=2000.0, sigma=1.0):
    """Smoothed value at one voxel as an explicit (2r+1)^3 sum (small grids only)."""
    src = np.nan_to_num(np.where(np.asarray(ct, np.float64) > metal_threshold, np.nan, np.asarray(ct, np.float64)))
    w = gaussian_weights(sigma)
    r = (len(w) - 1) // 2
    acc = 0.0
    for a in range(-r, r + 1):
        for b in range(-r, r + 1):
            for c in range(-r, r + 1):
                j = (reflect_index(idx[0] + a, src.shape[0]), reflect_index(idx[1] + b, src.shape[1]),
                     reflect_index(idx[2] + c, src.shape[2]))
                acc += w[a + r] * w[b + r] * w[c + r] * src[j]
    return acc


def density_from_hu(hu, materials):
    """Piecewise-constant HU -> density: material i covers [lo_i, lo_{i+1}),
    material 0 also everything below; NaN stays NaN."""
    hu = np.asarray(hu, dtype=np.float64)
    lo = np.array([m[0] for m in materials], dtype=np.float64)
    rho = np.array([m[2] for m in materials], dtype=np.float64)
    idx = np.clip(np.searchsorted(lo, hu, side="right") - 1, 0, len(lo) - 1)
    out = rho[idx]
    out[np.isnan(hu)] = np.nan
    return out
