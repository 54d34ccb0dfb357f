"""Drop-in dose calculators (reference layer L3/L4: ``pyvoxeldosimetry/core``)."""
from .activity_sampler import ActivitySampler
from .dose_calculator import DoseCalculationResult, DoseCalculator
from .dosimetry_base import DosimetryCalculator
from .kernel_convolution import KernelConvolutionCalculator
from .utils import (get_file_metadata, load_dose_map, load_kernel, resample_dose_map, save_dose_map, save_kernel,
                    verify_file_integrity)

__all__ = ["ActivitySampler", "DoseCalculationResult", "DoseCalculator", "DosimetryCalculator",
           "KernelConvolutionCalculator", "get_file_metadata", "load_dose_map", "load_kernel", "resample_dose_map",
           "save_dose_map", "save_kernel", "verify_file_integrity"]
