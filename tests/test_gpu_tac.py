"""GPU tests of the voxelwise TAC fits (include/pvd_tac.h) against oracle/tac.py
(scipy.optimize.curve_fit per voxel, as the reference).

Tolerances: both solvers converge to the weighted least-squares minimum with
MINPACK's default relative tolerance 1.49e-8; the GPU stores float32, so
well-conditioned parameters agree to ~1e-5 relative.
For models whose minimum is not unique on the sampled times (f3, bi-exponential),
the test checks that the GPU reaches a cost no higher than scipy's on at least
95 % of the voxels (a different descent path may end in another local minimum)."""
import sys

import numpy as np
import pytest

from oracle import tac

pytestmark = pytest.mark.gpu
T4 = [4.0, 24.0, 72.0, 144.0]
T6 = [1.0, 4.0, 24.0, 48.0, 96.0, 168.0]


def _say(*a):
    print("[gpu-case]", *a, file=sys.stderr, flush=True)


def _maps(y, shape):
    return [np.ascontiguousarray(y[:, i].reshape(shape), dtype=np.float32) for i in range(y.shape[1])]


def test_basic_fit_matches_curve_fit():
    from pyvoxeldosimetry_amd.time_integration import TimeCurveFitting
    _say("tac basic")
    rng = np.random.default_rng(0)
    shape = (8, 9, 10)
    n = int(np.prod(shape))
    A0 = rng.uniform(1e3, 1e5, n)
    lam = rng.uniform(0.004, 0.06, n)
    y = A0[:, None] * np.exp(-lam[:, None] * np.array(T4)) * rng.normal(1.0, 0.03, (n, 4))
    y[0] = 0.0                                   # all-zero voxel: p0 = (0, lambda), accumulated 0
    y[1] = [5.0, 10.0, 20.0, 40.0]               # growing curve: negative rate
    y = y.astype(np.float32).astype(np.float64)  # the GPU reads float32 maps
    w = [1.0, 2.0, 1.0, 0.5]
    tcf = TimeCurveFitting(64.1)
    p, acc = tcf.fit_time_activity_curve(T4, _maps(y, shape), weight_factors=w)
    assert p.shape == (2,) + shape and acc.shape == shape and p.dtype == np.float64
    pe, acce = tac.fit_basic(T4, y, 64.1, weights=w)
    np.testing.assert_allclose(p.reshape(2, -1), pe, rtol=2e-5, atol=1e-9)
    np.testing.assert_allclose(acc.reshape(-1), acce, rtol=5e-5)
    assert p.reshape(2, -1)[1, 0] == pytest.approx(np.log(2) / 64.1, rel=1e-6) and acc.reshape(-1)[0] == 0.0


@pytest.mark.parametrize("model,name", [("nukfit", "f2"), ("nukfit", "f3a"), ("lmfit", "mono-exponential"),
                                        ("nukfit", "f4a"), ("lmfit", "bi-exponential")])
def test_advanced_fit_reaches_the_least_squares_minimum(model, name):
    from pyvoxeldosimetry_amd.time_integration import TimeCurveFitting
    _say("tac advanced", model, name)
    rng = np.random.default_rng(hash(name) % 1000)
    hl = 64.1
    lp = np.log(2) / hl
    nuk, lmf = tac.models(hl, lp)
    f, p0 = (nuk if model == "nukfit" else lmf)[name]
    n = 120
    truth = {"f2": lambda: [rng.uniform(1, 5), rng.uniform(0.005, 0.05)],
             "f3a": lambda: [rng.uniform(1, 5), rng.uniform(0.2, 0.8), rng.uniform(0.02, 0.2)],
             "mono-exponential": lambda: [rng.uniform(1, 5), rng.uniform(20, 120)],
             "f4a": lambda: [rng.uniform(1, 3), rng.uniform(0.05, 0.2), rng.uniform(0.5, 1.5), rng.uniform(0.0, 0.01)],
             "bi-exponential": lambda: [rng.uniform(1, 5), rng.uniform(60, 120), rng.uniform(2, 10)]}[name]
    y = np.array([f(np.array(T6), *truth()) for _ in range(n)]) * rng.normal(1.0, 0.02, (n, len(T6)))
    y = y.astype(np.float32).astype(np.float64)
    shape = (4, 5, 6)
    tcf = TimeCurveFitting(hl)
    p, auc, info = tcf.fit_time_activity_curve_advanced(T6, _maps(y, shape), model=model, model_name=name,
                                                         return_fit_params=True)
    pe, auce, oke = tac.fit_advanced(T6, y, hl, model, name)
    pg = p.reshape(len(p0), -1)
    okg = np.array(info["success"])
    assert info["model_name"] == name and len(info["params"]) == n
    both = okg & oke
    assert okg.mean() >= 0.95 and both.mean() >= 0.9
    cg = np.array([tac.weighted_cost(f, T6, y[i], pg[:, i]) for i in range(n)])
    ce = np.array([tac.weighted_cost(f, T6, y[i], pe[:, i]) for i in range(n)])
    scale = np.array([np.sum(yy ** 2) for yy in y])
    as_good = cg[both] <= ce[both] * (1 + 1e-4) + 1e-9 * scale[both]
    if name in ("f2", "mono-exponential"):       # a unique minimum: the same parameters
        assert as_good.all()
        np.testing.assert_allclose(pg[:, both], pe[:, both], rtol=1e-4, atol=1e-7)
        np.testing.assert_allclose(auc.reshape(-1)[both], auce[both], rtol=1e-4)
    else:                                         # several local minima: a minimum at least as good, nearly always
        _say(f"{name}: GPU cost <= scipy cost on {as_good.mean():.3f} of {both.sum()} voxels")
        assert as_good.mean() >= 0.95
    # the integral is the trapezoid of the returned fit
    for i in np.nonzero(okg)[0][:20]:
        assert auc.reshape(-1)[i] == pytest.approx(tac.trapezoid(f(np.array(T6), *pg[:, i]), T6), rel=1e-5)


def test_more_parameters_than_samples_fail_like_the_reference():
    from pyvoxeldosimetry_amd.time_integration import TimeCurveFitting
    _say("tac underdetermined")
    y = np.array([[10.0, 8.0, 5.0, 3.0]] * 64)
    p, auc, info = TimeCurveFitting(64.1).fit_time_activity_curve_advanced(
        T4, _maps(y, (4, 4, 4)), "nukfit", "f4b", return_fit_params=True)
    assert not p.any() and not auc.any() and not any(info["success"])
    with pytest.raises(ValueError):
        TimeCurveFitting(64.1).fit_time_activity_curve_advanced(T4, _maps(y, (4, 4, 4)), model="other")
