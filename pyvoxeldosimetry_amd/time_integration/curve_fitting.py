"""Drop-in ``TimeCurveFitting`` (pyvoxeldosimetry/time_integration/curve_fitting.py:14-220)
over the HIP kernel of ``include/pvd_tac.h``.

The reference loops over voxels in Python and calls scipy.optimize.curve_fit
for each; here one GPU thread fits one voxel (FP64 Levenberg-Marquardt with
analytic Jacobians) and integrates it in the same pass.  Names, arguments,
defaults, start points and failure values follow the reference; outputs are
float64 NumPy arrays holding the float32 values the GPU writes.
"""
from __future__ import annotations

import ctypes
import math
from typing import List, Optional

import numpy as np

from .. import _native as nat

# name -> (kernel model id, start point); start points of curve_fitting.py:171-183
NUKFIT_MODELS = {
    "f2": (1, [1.0, 0.1]),
    "f3": (2, [1.0, 0.1, 0.01]),
    "f3a": (3, [1.0, 0.5, 0.1]),
    "f4a": (4, [1.0, 0.1, 0.5, 0.01]),
    "f4b": (5, [1.0, 0.1, 0.5, 0.01, 0.693]),
    "f4c": (6, [1.0, 0.1, 0.5, 0.01, 0.693]),
}
LMFIT_MODELS = ("mono-exponential", "bi-exponential", "tri-exponential")
PVD_TAC_MONO = 0
# Iteration caps mirror curve_fit's evaluation budgets (one LM iteration ~ n + 1
# evaluations with MINPACK's finite-difference Jacobian): default maxfev = 200 (n + 1)
# for the basic fit, maxfev = 10000 for the advanced one.  The tolerance is
# MINPACK's default ftol = xtol = 1.49012e-8 that curve_fit uses.
BASIC_MAX_ITER = 200
ADVANCED_MAXFEV = 10000
TOL = 1.49012e-8


def _lmfit_models(half_life):
    return {"mono-exponential": (7, [1.0, half_life]),
            "bi-exponential": (8, [1.0, half_life, half_life / 2]),
            "tri-exponential": (9, [1.0, half_life, 0.5, half_life / 2, half_life])}


class TimeCurveFitting:
    def __init__(self, half_life: float, device=None):
        self.half_life = half_life
        self.decay_constant = np.log(2) / half_life
        self.device = device

    # ------------------------------------------------------------------ GPU call
    def _fit(self, times, activities, model: int, p0, weight_factors, force_zero, max_iter, fail, auc_mode,
             t_int=0.0, lambda_phys=0.0, p0_from_data=False):
        import torch
        nat.require_device()
        lib = nat.load()
        times = np.asarray(times, dtype=np.float64)
        T = len(times)
        if not 1 <= T <= 16:
            raise ValueError("between 1 and 16 time points are supported")
        if len(activities) != T:
            raise ValueError("one activity map per time point")
        w = np.ones(T) if weight_factors is None else np.asarray(weight_factors, dtype=np.float64)
        dev = torch.device(self.device) if self.device is not None else torch.device("cuda",
                                                                                    torch.cuda.current_device())
        maps = []
        shape = None
        for a in activities:
            t = a if isinstance(a, torch.Tensor) else torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32))
            t = t.to(device=dev, dtype=torch.float32).contiguous()
            if shape is None:
                shape = tuple(t.shape)
            elif tuple(t.shape) != shape:
                raise ValueError("activity maps must share one shape")
            maps.append(t)
        n = int(np.prod(shape))
        q = nat.TacParams()
        q.model = model
        q.n_times = T
        for i in range(T):
            q.t[i] = times[i]
            q.w[i] = w[i]
        q.lambda_phys = lambda_phys
        for k, v in enumerate(p0):
            q.p0[k] = v
        q.p0_from_data = 1 if p0_from_data else 0
        q.force_zero_at_t0 = 1 if force_zero else 0
        q.max_iter = max_iter
        q.tol = TOL
        for k, v in enumerate(fail):
            q.fail_params[k] = v
        q.auc_mode = auc_mode
        q.t_int = t_int
        np_ = nat.check(lib.pvd_tac_n_params(model), "pvd_tac_n_params")
        params = torch.empty((np_, n), dtype=torch.float32, device=dev)
        auc = torch.empty(n, dtype=torch.float32, device=dev)
        status = torch.empty(n, dtype=torch.uint8, device=dev)
        arr = (ctypes.c_void_p * T)(*[m.data_ptr() for m in maps])
        with torch.cuda.device(dev):
            nat.check(lib.pvd_tac_fit(ctypes.byref(q), arr, n, ctypes.c_void_p(params.data_ptr()),
                                      ctypes.c_void_p(auc.data_ptr()), ctypes.c_void_p(status.data_ptr()),
                                      ctypes.c_void_p(nat.current_stream_handle(dev))), "pvd_tac_fit")
        return params, auc, status, shape

    # ------------------------------------------------------------------ reference API
    def fit_time_activity_curve(self, times: List[float], activities: List[np.ndarray],
                                weight_factors: Optional[List[float]] = None):
        """curve_fitting.py:19-65: A0 exp(-lam t) per voxel from p0 = (A(t_0), lambda_phys);
        failures give (0, lambda_phys).  Returns (params (2, *shape), accumulated (shape))
        with accumulated = A0/lam (1 - exp(-lam 100 T_half))."""
        params, auc, _, shape = self._fit(times, activities, PVD_TAC_MONO, [0.0, self.decay_constant],
                                          weight_factors, False, BASIC_MAX_ITER, [0.0, self.decay_constant], 0,
                                          t_int=100 * self.half_life, p0_from_data=True)
        return (params.cpu().numpy().astype(np.float64).reshape((2, *shape)),
                auc.cpu().numpy().astype(np.float64).reshape(shape))

    def _decay_function(self, t, A0, lambda_):
        """curve_fitting.py:67-72."""
        return A0 * np.exp(-lambda_ * t)

    def _calculate_accumulated_dose(self, fitted_params, integration_limit: float = None):
        """curve_fitting.py:74-84 (host arithmetic on already-fitted parameters)."""
        if integration_limit is None:
            integration_limit = 100 * self.half_life
        A0 = fitted_params[0]
        lambda_ = fitted_params[1]
        return A0 / lambda_ * (1 - np.exp(-lambda_ * integration_limit))

    @staticmethod
    def load_timepoint_images(image_paths_or_dirs):
        """curve_fitting.py:86-110 reads DICOM/NIfTI series; file formats are outside this package."""
        raise NotImplementedError("DICOM/NIfTI loading is not part of the GPU dose path; pass arrays")

    def fit_time_activity_curve_advanced(self, times: List[float], activities: List[np.ndarray],
                                         model: str = "nukfit", model_name: str = None,
                                         weight_factors: Optional[List[float]] = None,
                                         force_zero_at_t0: bool = False, integration_limit: float = None,
                                         phys_half_life: float = None, return_fit_params: bool = False):
        """curve_fitting.py:112-220: NUKFIT f2..f4c or LMFIT mono/bi/tri per voxel, then
        the trapezoid of the fitted curve over the sample times.  Failed voxels get zero
        parameters.  ``integration_limit`` is accepted and unused, as in the reference."""
        if model.lower() == "nukfit":
            models = dict(NUKFIT_MODELS)
        elif model.lower() == "lmfit":
            models = _lmfit_models(self.half_life)
        else:
            raise ValueError(f"Unknown model type: {model}")
        if model_name is None:
            model_name = list(models.keys())[0]
        mid, p0 = models[model_name]
        lam = self.decay_constant if phys_half_life is None else np.log(2) / phys_half_life
        params, auc, status, shape = self._fit(times, activities, mid, p0, weight_factors, force_zero_at_t0,
                                               ADVANCED_MAXFEV // (len(p0) + 1), [0.0] * len(p0), 1,
                                               lambda_phys=float(lam))
        fit_params = params.cpu().numpy().astype(np.float64).reshape((len(p0), *shape))
        aucs = auc.cpu().numpy().astype(np.float64).reshape(shape)
        if return_fit_params:
            flat = fit_params.reshape(len(p0), -1)
            fit_info = {"model": model, "model_name": model_name, "params": flat.T.tolist(),
                        "success": [bool(s) for s in status.cpu().numpy().tolist()]}
            return fit_params, aucs, fit_info
        return fit_params, aucs
