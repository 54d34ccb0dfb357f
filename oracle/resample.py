"""CPU oracle of the grid resampler (test infrastructure only).

The semantics are those declared in include/pvd_resample.h (SimpleITK's
ResampleImageFilter with origin 0, identity direction and default value 0,
which core/image_registration.py:128-173 uses; SimpleITK itself is not
installed here, so the restatement is pinned by closed forms instead):
continuous index c = j * out / in per axis, 0 outside [-0.5, n - 0.5),
trilinear with clamped neighbours (scipy.ndimage.map_coordinates order 1,
mode 'nearest') or nearest with ties rounded up.  FP64.
Pinned in tests/test_oracle_resample.py by: identity at equal spacing, exact
reproduction of affine functions inside the grid, an explicit per-voxel loop on
small grids, and constant fields (value inside, 0 outside).
"""
from __future__ import annotations

import numpy as np
from scipy.ndimage import map_coordinates


def output_shape(shape, sin, sout):
    return tuple(max(1, int(round(n * i / o))) for n, i, o in zip(shape, sin, sout))


def resample(img, sin, sout, interpolation="linear", shape=None, scale=1.0):
    img = np.asarray(img, np.float64)
    shape = tuple(shape) if shape is not None else output_shape(img.shape, sin, sout)
    grids = np.meshgrid(*[np.arange(m) * (o / i) for m, i, o in zip(shape, sin, sout)], indexing="ij")
    inside = np.ones(shape, bool)
    for g, n in zip(grids, img.shape):
        inside &= (g >= -0.5) & (g < n - 0.5)
    if interpolation == "nearest":
        idx = [np.clip(np.floor(g + 0.5).astype(np.int64), 0, n - 1) for g, n in zip(grids, img.shape)]
        out = img[tuple(idx)]
    else:
        out = map_coordinates(img, np.stack(grids), order=1, mode="nearest", prefilter=False)
    return np.where(inside, out, 0.0) * scale


def trilinear_at(img, c):
    """Explicit trilinear value at continuous index c with clamped neighbours (small grids)."""
    img = np.asarray(img, np.float64)
    acc = 0.0
    base = [int(np.floor(x)) for x in c]
    frac = [x - b for x, b in zip(c, base)]
    for dx in (0, 1):
        for dy in (0, 1):
            for dz in (0, 1):
                w = ((frac[0] if dx else 1 - frac[0]) * (frac[1] if dy else 1 - frac[1])
                     * (frac[2] if dz else 1 - frac[2]))
                idx = tuple(min(max(b + d, 0), n - 1) for b, d, n in zip(base, (dx, dy, dz), img.shape))
                acc += w * img[idx]
    return acc
