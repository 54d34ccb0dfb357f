"""Hypothesis property tests on the CPU (bounded, derandomised examples).

The convolution oracle (oracle/conv.py, the restatement of
kernel_convolution.py:71-74 and its 'same' linear variant) over random shapes:
linearity, circular shift equivariance, a delta giving back the (cropped or
padded) kernel, and agreement with brute-force sums on tiny grids.  The
trapezoid weights (kernel_convolution.py:78-106) integrate piecewise-linear
activity curves exactly.

The host-side plan layout (pvd_plan_describe, no device call) over random
volume and kernel shapes: FFT lengths that never alias the reference's circular
result or the linear result, 7-smooth lengths, the 7-smooth rule restored by
PVD_FLAG_SMOOTH_LENGTHS, register paths only where their lengths are, and
launch shapes within the two-step instances' block bound.
"""
import numpy as np
import pytest
from hypothesis import HealthCheck, given, settings
from hypothesis import strategies as st

from oracle import conv

SETTINGS = dict(max_examples=40, deadline=None, derandomize=True,
                suppress_health_check=[HealthCheck.too_slow])

dims = st.tuples(st.integers(1, 7), st.integers(1, 7), st.integers(1, 7))


def _arr(seed, shape):
    return np.random.default_rng(seed).normal(size=shape)


# --------------------------------------------------------------------- oracle
@settings(**SETTINGS)
@given(n=dims, k=dims, seed=st.integers(0, 2 ** 31 - 1), a=st.floats(-3, 3), b=st.floats(-3, 3))
def test_reference_convolution_is_linear_in_the_activity(n, k, seed, a, b):
    x, y, kern = _arr(seed, n), _arr(seed + 1, n), _arr(seed + 2, k)
    lhs = conv.ref_conv(a * x + b * y, kern)
    rhs = a * conv.ref_conv(x, kern) + b * conv.ref_conv(y, kern)
    np.testing.assert_allclose(lhs, rhs, rtol=1e-9, atol=1e-9 * (1 + np.abs(rhs).max()))


@settings(**SETTINGS)
@given(n=dims, k=dims, seed=st.integers(0, 2 ** 31 - 1), s=st.tuples(*[st.integers(-9, 9)] * 3))
def test_reference_convolution_commutes_with_circular_shifts(n, k, seed, s):
    x, kern = _arr(seed, n), _arr(seed + 1, k)
    lhs = conv.ref_conv(np.roll(x, s, axis=(0, 1, 2)), kern)
    rhs = np.roll(conv.ref_conv(x, kern), s, axis=(0, 1, 2))
    np.testing.assert_allclose(lhs, rhs, rtol=1e-9, atol=1e-9 * (1 + np.abs(rhs).max()))


@settings(**SETTINGS)
@given(n=dims, k=dims, seed=st.integers(0, 2 ** 31 - 1), data=st.data())
def test_delta_gives_back_the_padded_or_cropped_kernel_rolled_to_the_source(n, k, seed, data):
    p = tuple(data.draw(st.integers(0, m - 1)) for m in n)
    kern = _arr(seed, k)
    x = np.zeros(n)
    x[p] = 1.0
    # the kernel cut or zero-padded to the grid at its high end, anchored at index 0
    kp = np.zeros(n)
    sl = tuple(slice(0, min(a, b)) for a, b in zip(n, k))
    kp[sl] = kern[sl]
    np.testing.assert_allclose(conv.ref_conv(x, kern), np.roll(kp, p, axis=(0, 1, 2)), rtol=0, atol=1e-12)


@settings(**SETTINGS)
@given(n=dims, k=dims, seed=st.integers(0, 2 ** 31 - 1))
def test_reference_and_linear_convolutions_match_brute_force_sums(n, k, seed):
    x, kern = _arr(seed, n), _arr(seed + 1, k)
    np.testing.assert_allclose(conv.ref_conv(x, kern), conv.brute_circular(x, kern), rtol=1e-9, atol=1e-10)
    np.testing.assert_allclose(conv.linear_same(x, kern), conv.brute_linear_same(x, kern), rtol=1e-9, atol=1e-10)


@settings(**SETTINGS)
@given(n=dims, k=dims, seed=st.integers(0, 2 ** 31 - 1), data=st.data())
def test_linear_delta_is_the_centred_kernel_cut_at_the_grid(n, k, seed, data):
    p = tuple(data.draw(st.integers(0, m - 1)) for m in n)
    kern = _arr(seed, k)
    x = np.zeros(n)
    x[p] = 1.0
    exp = np.zeros(n)
    for i in np.ndindex(*n):
        j = tuple(ii - pp + kk // 2 for ii, pp, kk in zip(i, p, k))
        if all(0 <= jj < kk for jj, kk in zip(j, k)):
            exp[i] = kern[j]
    np.testing.assert_allclose(conv.linear_same(x, kern), exp, rtol=0, atol=1e-12)


@settings(**SETTINGS)
@given(t=st.lists(st.floats(0, 500, allow_nan=False), min_size=2, max_size=8, unique=True),
       seed=st.integers(0, 2 ** 31 - 1))
def test_trapezoid_weights_integrate_piecewise_linear_curves_exactly(t, seed):
    t = np.sort(np.asarray(t))
    if np.min(np.diff(t)) < 1e-3:
        return
    a = np.random.default_rng(seed).uniform(0, 10, size=t.size)
    w = conv.trapezoid_weights_seconds(t)
    exact = sum(0.5 * (a[i] + a[i + 1]) * (t[i + 1] - t[i]) for i in range(t.size - 1)) * 3600.0
    assert np.dot(w, a) == pytest.approx(exact, rel=1e-12, abs=1e-9)


# --------------------------------------------------------------------- host plan layout
def _smooth(n):
    for p in (2, 3, 5, 7):
        while n % p == 0:
            n //= p
    return n == 1


# register transforms: 256 = 16 x 16 on y/x, P = 16 B on y/x, P = A B on z
XY_TWO_STEP = {16 * b for b in (6, 8, 9, 10, 12, 14, 15, 16, 18, 20, 21, 24, 25, 27, 28, 30, 32, 35, 36)}
Z_TWO_STEP = {8 * b for b in (16, 20, 24, 25, 30, 12, 14, 15, 18, 21, 27)} | {16 * b for b in (16, 20, 25, 27, 14, 18, 21,
                                                                                             24, 30)}


def _path_fits(axis, path, p):
    if path == "sq256":
        return axis < 2 and p == 256
    if path == "two_step":
        return p in (Z_TWO_STEP if axis == 2 else XY_TWO_STEP)
    return path == "stockham"


@settings(max_examples=150, deadline=None, derandomize=True)
@given(shape=st.tuples(*[st.integers(1, 700)] * 3), k=st.tuples(*[st.integers(1, 61)] * 3),
       mode=st.sampled_from(["reference", "linear"]))
def test_plan_lengths_never_alias_and_paths_fit_their_lengths(shape, k, mode):
    from pyvoxeldosimetry_amd import describe_plan
    d = describe_plan(shape, k, mode, algo="fft")
    s = describe_plan(shape, k, mode, algo="fft", smooth_lengths=True)
    assert d["algo"] == "fft" and d["n_passes"] == 5
    for a in range(3):
        n, ka, p = shape[a], k[a], d["fft_n"][a]
        assert _smooth(p)
        if mode == "reference":
            # the circular result of size n: either exactly n, or long enough that no
            # product of the (cropped) kernel wraps before the fold back onto n
            need = n + min(ka, n) - 1
            assert p == n or p >= need, (a, n, ka, p)
        else:
            need = n + min(ka // 2, n - 1)
            assert p >= need, (a, n, ka, p)
        # taps used: the crop [0:n] (reference), the taps within n - 1 of the centre (linear)
        c = ka // 2
        used = min(ka, n) if mode == "reference" else min(c, n - 1) + min(ka - 1 - c, n - 1) + 1
        assert d["kernel_used"][a] == used
        # the 7-smooth rule alone: the smallest admissible 7-smooth length
        p0 = s["fft_n"][a]
        assert _smooth(p0) and p0 <= p <= 2 * p0
        if p != p0:
            assert p0 > 128 and d["fft_path"][a] != "stockham"
        path = d["fft_path"][a]
        assert _path_fits(a, path, p), (a, p, path)
    # the two-step y/x instances run blocks of at most 512 threads
    for a, pas in ((1, (1, 3)), (0, (2,))):
        if d["fft_path"][a] == "two_step":
            assert all(d["block"][i] <= 512 for i in pas)
    assert d["workspace_bytes"] >= 8 * shape[0] * d["fft_n"][1] * (d["fft_n"][2] // 2 + 1)


# --------------------------------------------------------------------- tiled network slabs
@settings(max_examples=200, deadline=None, derandomize=True)
@given(nz=st.integers(32, 2000), k=st.integers(1, 16))
def test_network_slabs_tile_the_axis_and_keep_the_receptive_field(nz, k):
    """Slabs of extent e (a multiple of 32) whose kept ranges tile [0, nz) exactly and lie
    at least 6 voxels from every cut face (the receptive field of six 3^3 layers)."""
    from pyvoxeldosimetry_amd.learned.pipeline import z_slabs
    e = min(32 * k, nz // 32 * 32)
    sl = z_slabs(nz, e)
    assert sl[0][2] == 0 and sl[-1][3] == nz
    for (_, _, _, b), (_, _, c, _) in zip(sl, sl[1:]):
        assert b == c
    for s0, s1, a, b in sl:
        assert s1 - s0 == e and 0 <= s0 <= a < b <= s1 <= nz
        assert a == 0 or a - s0 >= 6
        assert b == nz or s1 - b >= 6


@settings(max_examples=200, deadline=None, derandomize=True)
@given(nz=st.integers(32, 1200), plane=st.sampled_from([1, 64 * 64, 512 * 512]))
def test_automatic_slab_extent_computes_the_fewest_planes(nz, plane):
    """slab_extent: a multiple of 32 within [32, nz], within the memory cap (or 32), and
    no allowed extent computes fewer planes (brute force over every extent)."""
    from pyvoxeldosimetry_amd.learned.pipeline import slab_extent, z_slabs
    cap = 8 * 2 ** 30
    e = slab_extent(nz, plane, max_bytes=cap)
    assert e % 32 == 0 and 32 <= e <= nz
    per = plane * (16 + 5 * 32) * 2
    assert e == 32 or per * e <= cap
    planes = lambda x: sum(s1 - s0 for s0, s1, _, _ in z_slabs(nz, x))  # noqa: E731
    allowed = [x for x in range(32, nz // 32 * 32 + 1, 32) if x == 32 or per * x <= cap]
    assert planes(e) == min(planes(x) for x in allowed)
    if nz == 400 and plane == 512 * 512:
        assert e == 224 and planes(e) == 448
