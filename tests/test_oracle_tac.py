"""Pins of the TAC-fit oracle (oracle/tac.py) on the CPU."""
import numpy as np
import pytest

from oracle import tac

T = [4.0, 24.0, 72.0, 144.0]


def test_noise_free_mono_recovers_parameters():
    rng = np.random.default_rng(0)
    A0 = rng.uniform(1e3, 1e5, 20)
    lam = rng.uniform(0.005, 0.05, 20)
    y = A0[:, None] * np.exp(-lam[:, None] * np.array(T))
    p, acc = tac.fit_basic(T, y, 64.1)
    np.testing.assert_allclose(p[0], A0, rtol=1e-6)
    np.testing.assert_allclose(p[1], lam, rtol=1e-6)
    np.testing.assert_allclose(acc, A0 / lam * (1 - np.exp(-lam * 6410.0)), rtol=1e-6)


def test_two_samples_interpolate_exactly():
    t = [2.0, 30.0]
    y = np.array([[1000.0, 400.0], [50.0, 49.0]])
    p, _ = tac.fit_basic(t, y, 64.1)
    lam = np.log(y[:, 0] / y[:, 1]) / (t[1] - t[0])
    np.testing.assert_allclose(p[1], lam, rtol=1e-7)
    np.testing.assert_allclose(p[0], y[:, 0] * np.exp(lam * t[0]), rtol=1e-7)


def test_trapezoid_matches_numpy():
    rng = np.random.default_rng(1)
    x = np.sort(rng.uniform(0, 100, 7))
    y = rng.normal(size=7)
    assert tac.trapezoid(y, x) == pytest.approx(np.trapezoid(y, x), rel=1e-14)


def test_more_parameters_than_samples_fail_with_zeros():
    y = np.array([[10.0, 8.0, 5.0, 3.0]])
    p, auc, ok = tac.fit_advanced(T, y, 64.1, "nukfit", "f4b")      # 5 parameters, 4 samples
    assert not ok[0] and not p.any() and auc[0] == 0.0


def test_advanced_models_recover_noise_free_curves():
    hl = 64.1
    lp = np.log(2) / hl
    t = np.array([1.0, 4.0, 24.0, 48.0, 96.0, 168.0])
    nuk, lmf = tac.models(hl, lp)
    truth = {"f2": [3.0, 0.02], "f3a": [2.0, 0.6, 0.05], "mono-exponential": [5.0, 40.0]}
    for name, p_true in truth.items():
        f = (nuk.get(name) or lmf[name])[0]
        y = f(t, *p_true)[None, :]
        p, auc, ok = tac.fit_advanced(t, y, hl, "nukfit" if name in nuk else "lmfit", name)
        assert ok[0]
        np.testing.assert_allclose(p[:, 0], p_true, rtol=1e-5)
        assert auc[0] == pytest.approx(np.trapezoid(y[0], t), rel=1e-6)
