"""Pins of the resampling oracle (oracle/resample.py) on the CPU."""
import numpy as np
import pytest

from oracle import resample as ors


def test_equal_spacing_is_identity():
    rng = np.random.default_rng(0)
    img = rng.normal(size=(7, 5, 9))
    for mode in ("linear", "nearest"):
        np.testing.assert_array_equal(ors.resample(img, (2, 2, 2), (2, 2, 2), mode), img)


def test_affine_function_reproduced_inside_and_clamped_outside():
    n = (9, 8, 10)
    x, y, z = np.meshgrid(*[np.arange(k, dtype=float) for k in n], indexing="ij")
    img = 3.0 * x - 2.0 * y + 0.5 * z + 7.0
    out = ors.resample(img, (2.0, 2.0, 2.0), (1.3, 0.7, 1.9), "linear")
    c = np.meshgrid(*[np.arange(m) * (o / 2.0) for m, o in zip(out.shape, (1.3, 0.7, 1.9))], indexing="ij")
    cl = [np.clip(ci, 0, k - 1) for ci, k in zip(c, n)]
    inside = np.all([(ci >= -0.5) & (ci < k - 0.5) for ci, k in zip(c, n)], axis=0)
    exp = np.where(inside, 3.0 * cl[0] - 2.0 * cl[1] + 0.5 * cl[2] + 7.0, 0.0)
    np.testing.assert_allclose(out, exp, rtol=1e-12, atol=1e-12)
    assert out.shape == ors.output_shape(n, (2.0,) * 3, (1.3, 0.7, 1.9))


def test_matches_explicit_loop():
    rng = np.random.default_rng(1)
    img = rng.normal(size=(4, 3, 5))
    sin, sout = (1.0, 2.0, 1.5), (0.6, 1.1, 2.2)
    out = ors.resample(img, sin, sout, "linear")
    for idx in np.ndindex(out.shape):
        c = [j * o / i for j, i, o in zip(idx, sin, sout)]
        inside = all(-0.5 <= ci < k - 0.5 for ci, k in zip(c, img.shape))
        assert out[idx] == pytest.approx(ors.trilinear_at(img, c) if inside else 0.0, abs=1e-12)


def test_constant_field_and_nearest_ties():
    img = np.full((4, 4, 4), 2.5)
    out = ors.resample(img, (1.0,) * 3, (0.5,) * 3, "nearest")   # c = j/2: 3.5 is outside, 3.0 inside
    assert out.shape == (8, 8, 8)
    assert (out[:7, :7, :7] == 2.5).all() and (out[7] == 0).all()
    line = np.arange(4, dtype=float).reshape(4, 1, 1) * np.ones((1, 1, 1))
    got = ors.resample(line, (1.0,) * 3, (0.5, 1.0, 1.0), "nearest")[:, 0, 0]
    assert got.tolist() == [0, 1, 1, 2, 2, 3, 3, 0]            # c = 0.5 -> 1 (half up), 3.5 outside


def _coefficients_3d(img):
    c = np.asarray(img, np.float64)
    for ax in range(3):
        c = np.apply_along_axis(ors.bspline_coefficients_1d, ax, c)
    return c


def test_cubic_matches_the_explicit_mirror_boundary_spline():
    rng = np.random.default_rng(4)
    for shape in ((5, 4, 6), (1, 3, 7), (2, 6, 3)):
        img = rng.normal(size=shape)
        sin, sout = (1.0, 1.5, 2.0), (0.7, 1.1, 1.3)
        out = ors.resample(img, sin, sout, "cubic")
        coef = _coefficients_3d(img)
        for idx in np.ndindex(out.shape):
            c = [j * o / i for j, i, o in zip(idx, sin, sout)]
            inside = all(-0.5 <= ci < k - 0.5 for ci, k in zip(c, img.shape))
            assert out[idx] == pytest.approx(ors.bspline_at(coef, c) if inside else 0.0, abs=1e-9), (shape, idx)


def test_cubic_interpolates_samples_and_reproduces_cubics_away_from_the_edges():
    rng = np.random.default_rng(5)
    img = rng.normal(size=(6, 7, 5))
    np.testing.assert_allclose(ors.resample(img, (2.0,) * 3, (2.0,) * 3, "cubic"), img, rtol=0, atol=1e-12)
    n = (48, 44, 40)
    x, y, z = np.meshgrid(*[np.arange(k, dtype=float) for k in n], indexing="ij")
    f = lambda a, b, c: 0.01 * a ** 3 - 0.02 * b ** 2 * c + 0.5 * a * b - c + 3.0   # noqa: E731
    out = ors.resample(f(x, y, z), (1.0,) * 3, (0.55, 0.6, 0.7), "cubic")
    g = np.meshgrid(*[np.arange(m) * o for m, o in zip(out.shape, (0.55, 0.6, 0.7))], indexing="ij")
    interior = np.all([(gi >= 16) & (gi <= k - 17) for gi, k in zip(g, n)], axis=0)
    assert interior.sum() > 1000
    exp = f(*g)
    np.testing.assert_allclose(out[interior], exp[interior], rtol=0, atol=1e-6 * np.abs(exp).max())
