"""CPU oracle of the grid resampler (test infrastructure only).

The semantics are those declared in include/pvd_resample.h (SimpleITK's
ResampleImageFilter with origin 0, identity direction and default value 0,
which core/image_registration.py:128-173 uses; SimpleITK itself is not
installed here, so the restatement is pinned by closed forms instead):
continuous index c = j * out / in per axis, 0 outside [-0.5, n - 0.5),
trilinear with clamped neighbours (scipy.ndimage.map_coordinates order 1,
mode 'nearest'), nearest with ties rounded up, or cubic: the interpolating cubic
B-spline with mirror boundaries (SimpleITK's sitkBSpline; map_coordinates order 3,
mode 'mirror', prefilter on).  FP64.
Pinned in tests/test_oracle_resample.py by: identity at equal spacing, exact
reproduction of affine functions inside the grid, an explicit per-voxel loop on
small grids, and constant fields (value inside, 0 outside); cubic by the explicit
spline (coefficients from the tridiagonal mirror-boundary system, B-spline
weights), interpolation of the samples, and cubic polynomials reproduced away
from the edges.
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
    elif interpolation == "cubic":
        out = map_coordinates(img, np.stack(grids), order=3, mode="mirror", prefilter=True)
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


def bspline_coefficients_1d(f):
    """Coefficients c of the interpolating cubic B-spline with mirror boundaries:
    f[i] = (c[i-1] + 4 c[i] + c[i+1]) / 6 with c[-1] = c[1], c[n] = c[n-2] (dense solve)."""
    f = np.asarray(f, np.float64)
    n = f.shape[0]
    if n == 1:
        return f.copy()
    a = np.zeros((n, n))
    for i in range(n):
        a[i, i] += 4.0
        a[i, abs(i - 1) if i > 0 else 1] += 1.0
        a[i, i + 1 if i < n - 1 else n - 2] += 1.0
    return np.linalg.solve(a / 6.0, f)


def bspline_at(coef, c):
    """Explicit cubic B-spline value at continuous index c (mirror-indexed neighbours)."""
    def mirror(i, n):
        if n == 1:
            return 0
        p = 2 * (n - 1)
        m = i % p
        return m if m < n else p - m

    def w(t):
        return [(1 - t) ** 3 / 6, (3 * t ** 3 - 6 * t ** 2 + 4) / 6, (-3 * t ** 3 + 3 * t ** 2 + 3 * t + 1) / 6, t ** 3 / 6]

    base = [int(np.floor(x)) for x in c]
    ws = [w(x - b) for x, b in zip(c, base)]
    acc = 0.0
    for a in range(4):
        for b in range(4):
            for d in range(4):
                idx = tuple(mirror(bb - 1 + o, n) for bb, o, n in zip(base, (a, b, d), coef.shape))
                acc += ws[0][a] * ws[1][b] * ws[2][d] * coef[idx]
    return acc
