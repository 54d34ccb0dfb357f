"""Learned heterogeneity correction: Monte Carlo targets and a 3-D residual CNN."""
from .cnn import ResidualCNN
from .mc import MonteCarlo

__all__ = ["MonteCarlo", "ResidualCNN"]
