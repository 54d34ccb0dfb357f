"""Drop-in dose calculators (reference layer L3/L4: ``pyvoxeldosimetry/core``)."""
from .activity_sampler import ActivitySampler
from .dose_calculator import DoseCalculationResult, DoseCalculator
from .dosimetry_base import DosimetryCalculator
from .kernel_convolution import KernelConvolutionCalculator

__all__ = ["ActivitySampler", "DoseCalculationResult", "DoseCalculator", "DosimetryCalculator",
           "KernelConvolutionCalculator"]
