"""CPU oracle of the voxelwise TAC fits (test infrastructure only).

Restates pyvoxeldosimetry/time_integration/curve_fitting.py:19-84 and :112-220
with the same library call (scipy.optimize.curve_fit per voxel, pure-Python
loop, so small voxel counts only).  Pinned in tests/test_oracle_tac.py by
closed forms: noise-free exponentials recover their parameters, two samples
give the exact interpolating exponential, the tail integral and the trapezoid
match their formulas, and more parameters than samples fail as in the
reference.
"""
from __future__ import annotations

import warnings

import numpy as np
from scipy.optimize import curve_fit

LN2 = np.log(2)


def mono(t, A0, lam):
    return A0 * np.exp(-lam * t)


def models(half_life, lambda_phys):
    lp = lambda_phys
    nuk = {
        "f2": (lambda t, A1, l1: A1 * np.exp(-(l1 + lp) * t), [1.0, 0.1]),
        "f3": (lambda t, A1, l1, l2: A1 * (np.exp(-(l1 + lp) * t) - np.exp(-(l2 + lp) * t)), [1.0, 0.1, 0.01]),
        "f3a": (lambda t, A1, a, l1: A1 * (a * np.exp(-(l1 + lp) * t) + (1 - a) * np.exp(-lp * t)), [1.0, 0.5, 0.1]),
        "f4a": (lambda t, A1, l1, A2, l2: A1 * np.exp(-(l1 + lp) * t) + A2 * np.exp(-(l2 + lp) * t),
                [1.0, 0.1, 0.5, 0.01]),
        "f4b": (lambda t, A1, l1, A2, l2, lb: (A1 * np.exp(-(l1 + lp) * t) - A2 * np.exp(-(l2 + lp) * t)
                                               - (A1 - A2) * np.exp(-(lb + lp) * t)), [1.0, 0.1, 0.5, 0.01, 0.693]),
        "f4c": (lambda t, A1, l1, A2, l2, lb: (A1 * np.exp(-(l1 + lp) * t) + A2 * np.exp(-(l2 + lp) * t)
                                               - (A1 + A2) * np.exp(-(lb + lp) * t)), [1.0, 0.1, 0.5, 0.01, 0.693]),
    }
    lmf = {
        "mono-exponential": (lambda t, A1, T1: A1 * np.exp(-LN2 * t / T1), [1.0, half_life]),
        "bi-exponential": (lambda t, A1, T1, T2: A1 * (np.exp(-LN2 * t / T1) - np.exp(-LN2 * t / T2)),
                           [1.0, half_life, half_life / 2]),
        "tri-exponential": (lambda t, A1, T1, A2, T2, Tp: (A1 * np.exp(-LN2 * t / T1) - A2 * np.exp(-LN2 * t / T2)
                                                            - (A1 - A2) * np.exp(-LN2 * t / Tp)),
                            [1.0, half_life, 0.5, half_life / 2, half_life]),
    }
    return nuk, lmf


def fit_basic(times, y_voxels, half_life, weights=None):
    """curve_fitting.py:19-65 for an [n_voxels, T] array: (params [2, n], accumulated [n])."""
    times = np.asarray(times, np.float64)
    lam0 = np.log(2) / half_life
    w = np.ones_like(times) if weights is None else np.asarray(weights, np.float64)
    out = np.zeros((2, len(y_voxels)))
    for i, y in enumerate(np.asarray(y_voxels, np.float64)):
        try:
            with warnings.catch_warnings():
                warnings.simplefilter("ignore")
                popt, _ = curve_fit(mono, times, y, p0=[y[0], lam0], sigma=1 / w)
            out[:, i] = popt
        except Exception:
            out[:, i] = [0, lam0]
    acc = accumulated(out, half_life)
    return out, acc


def accumulated(params, half_life, integration_limit=None):
    """curve_fitting.py:74-84."""
    if integration_limit is None:
        integration_limit = 100 * half_life
    with np.errstate(divide="ignore", invalid="ignore"):
        return params[0] / params[1] * (1 - np.exp(-params[1] * integration_limit))


def fit_advanced(times, y_voxels, half_life, model="nukfit", model_name=None, weights=None, force_zero=False,
                 phys_half_life=None):
    """curve_fitting.py:112-220 for an [n_voxels, T] array: (params [P, n], aucs [n], success [n])."""
    times = np.asarray(times, np.float64)
    lp = np.log(2) / half_life if phys_half_life is None else np.log(2) / phys_half_life
    nuk, lmf = models(half_life, lp)
    table = nuk if model.lower() == "nukfit" else lmf
    if model_name is None:
        model_name = list(table)[0]
    f, p0 = table[model_name]
    w = np.ones_like(times) if weights is None else np.asarray(weights, np.float64)
    n = len(y_voxels)
    params = np.zeros((len(p0), n))
    aucs = np.zeros(n)
    ok = np.zeros(n, bool)
    for i, y in enumerate(np.asarray(y_voxels, np.float64)):
        y = y.copy()
        if force_zero:
            y = y - y[0]
            y[0] = 0
        try:
            with warnings.catch_warnings():
                warnings.simplefilter("ignore")
                popt, _ = curve_fit(f, times, y, p0=p0, sigma=1 / w, maxfev=10000)
            params[:, i] = popt
            ok[i] = True
        except Exception:
            params[:, i] = 0
        with np.errstate(all="ignore"):
            aucs[i] = trapezoid(f(times, *params[:, i]), times)
    return params, aucs, ok


def trapezoid(y, x):
    """np.trapz(y, x) (the reference's call), written out."""
    y = np.asarray(y, np.float64)
    x = np.asarray(x, np.float64)
    return float(np.sum((x[1:] - x[:-1]) * (y[1:] + y[:-1]) / 2.0))


def weighted_cost(f, times, y, params, weights=None):
    """sum (w (y - f))^2 of one voxel."""
    w = np.ones(len(times)) if weights is None else np.asarray(weights, np.float64)
    with np.errstate(all="ignore"):
        return float(np.sum((w * (np.asarray(y, np.float64) - f(np.asarray(times, np.float64), *params))) ** 2))
