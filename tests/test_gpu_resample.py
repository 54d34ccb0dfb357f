"""GPU tests of the grid resampler (include/pvd_resample.h) against oracle/resample.py,
and of the dose path with a voxel size other than the kernel resolution.

Tolerances: FP64 coordinates on both sides, FP32 interpolation on the GPU:
|diff| <= 1e-6 x max|image|.  Nearest is exact."""
import sys

import numpy as np
import pytest

from oracle import conv, dpk
from oracle import resample as ors

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[gpu-case]", *a, file=sys.stderr, flush=True)


@pytest.mark.parametrize("sin,sout", [((2.0, 2.0, 2.0), (1.0, 1.0, 1.0)), ((1.0, 1.0, 1.0), (2.2, 2.2, 2.2)),
                                      ((0.98, 0.98, 2.5), (2.0, 1.3, 1.7)), ((3.0, 3.0, 3.0), (3.0, 3.0, 3.0))])
def test_resample_matches_oracle(sin, sout):
    from pyvoxeldosimetry_amd.resample import resample
    _say("resample", sin, sout)
    rng = np.random.default_rng(3)
    img = rng.normal(size=(33, 27, 41)).astype(np.float32)
    for mode in ("linear", "nearest"):
        got = resample(img, sin, sout, mode)
        exp = ors.resample(img, sin, sout, mode)
        assert got.shape == exp.shape
        if mode == "nearest":
            np.testing.assert_array_equal(got, exp.astype(np.float32))
        else:
            np.testing.assert_allclose(got, exp, rtol=0, atol=1e-6 * np.abs(img).max())


def test_activity_resampling_keeps_total():
    from pyvoxeldosimetry_amd.phantoms import create_spherical_activity
    from pyvoxeldosimetry_amd.resample import resample_activity
    _say("resample activity total")
    act = create_spherical_activity((64, 64, 64), (32, 32, 32), 12, 1000.0)
    from scipy.ndimage import gaussian_filter
    act = gaussian_filter(act, 2.0).astype(np.float32)            # smooth: interpolation keeps the integral
    down = resample_activity(act, (1.0,) * 3, (2.0,) * 3)
    up = resample_activity(act, (1.0,) * 3, (0.5,) * 3)
    assert down.sum() == pytest.approx(act.sum(), rel=2e-3)
    assert up.sum() == pytest.approx(act.sum(), rel=2e-3)


def test_dose_at_other_voxel_size_goes_through_the_kernel_grid():
    """2 mm activity, 1 mm Y-90 kernel: activity and density onto the 1 mm grid,
    convolution there, dose back onto the 2 mm grid (the reference's intended path)."""
    from pyvoxeldosimetry_amd import KernelConvolutionCalculator
    _say("dose with resampling")
    rng = np.random.default_rng(5)
    act = np.zeros((24, 20, 22), np.float32)
    act[8:14, 6:13, 9:15] = rng.uniform(1e4, 2e4, size=(6, 7, 6))
    rho = rng.uniform(0.8, 1.2, size=act.shape).astype(np.float32)
    calc = KernelConvolutionCalculator("Y90", "water", 1.0, {"kernel_grid": (15, 15, 15), "fix_kernel_centre": True,
                                                             "conv_mode": "linear"})
    got = calc.calculate_dose_rate(act, (2.0, 2.0, 2.0), density_map=rho)
    ak = ors.resample(act, (2.0,) * 3, (1.0,) * 3, "linear", scale=1.0 / 8.0)
    rk = ors.resample(rho, (2.0,) * 3, (1.0,) * 3, "linear", shape=ak.shape)
    dk = conv.epilogue(conv.linear_same(ak, dpk.y90_kernel(1.0, (15, 15, 15), fix_centre=True)), rk)
    exp = ors.resample(dk, (1.0,) * 3, (2.0,) * 3, "linear", shape=act.shape)
    assert got.shape == act.shape
    assert conv.error_metrics(got, exp)["max_rel"] < 1e-5


@pytest.mark.parametrize("sin,sout,shape", [((2.0, 2.0, 2.0), (0.98, 0.98, 0.98), (19, 23, 17)),
                                            ((0.98, 1.2, 2.5), (2.0, 1.0, 1.7), (41, 30, 22)),
                                            ((1.0, 1.0, 1.0), (1.0, 1.0, 1.0), (5, 1, 9)),
                                            ((1.0, 1.0, 1.0), (0.3, 0.4, 0.35), (2, 3, 4))])
def test_cubic_bspline_matches_oracle(sin, sout, shape):
    """interpolation='cubic' (sitkBSpline): mirror-boundary interpolating cubic B-spline;
    FP32 prefilter and evaluation against the FP64 oracle within 2e-6 of max|image|."""
    from pyvoxeldosimetry_amd.resample import resample
    _say("cubic", sin, sout, shape)
    rng = np.random.default_rng(sum(shape))
    img = rng.normal(size=shape).astype(np.float32)
    got = resample(img, sin, sout, "cubic")
    exp = ors.resample(img, sin, sout, "cubic")
    assert got.shape == exp.shape
    np.testing.assert_allclose(got, exp, rtol=0, atol=2e-6 * np.abs(img).max())


def test_resample_dose_map_shape_affine_and_values():
    from pyvoxeldosimetry_amd.core import resample_dose_map
    _say("resample_dose_map")
    rng = np.random.default_rng(9)
    dose = rng.random((30, 26, 21)).astype(np.float32)
    out, aff = resample_dose_map(dose, (2.0, 2.0, 2.5), (0.98, 1.5, 2.0))
    shape = tuple(int(np.ceil(n * i / o)) for n, i, o in zip(dose.shape, (2.0, 2.0, 2.5), (0.98, 1.5, 2.0)))
    assert out.shape == shape
    np.testing.assert_array_equal(aff, np.diag([0.98, 1.5, 2.0, 1.0]))
    exp = ors.resample(dose, (2.0, 2.0, 2.5), (0.98, 1.5, 2.0), "cubic", shape=shape)
    np.testing.assert_allclose(out, exp, rtol=0, atol=2e-6)
