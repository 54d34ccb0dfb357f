"""Voxelwise time-activity-curve fitting and integration on the GPU."""
from .curve_fitting import LMFIT_MODELS, NUKFIT_MODELS, TimeCurveFitting

__all__ = ["LMFIT_MODELS", "NUKFIT_MODELS", "TimeCurveFitting"]
