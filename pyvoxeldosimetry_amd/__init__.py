"""MI355X-native kernel-convolution dose path of PyVoxelDosimetry.

Python host code mirroring the reference's plugin API over hand-written HIP
kernels for gfx950 (``libpvd.so``, C ABI in ``include/pvd.h``).
"""
from .core import (ActivitySampler, DoseCalculationResult, DoseCalculator, DosimetryCalculator,
                   KernelConvolutionCalculator)
from .dose_kernels import Ga68KernelGenerator, KernelFactory, Lu177KernelGenerator, Y90KernelGenerator
from .plan import ConvPlan, describe_plan

__version__ = "0.1.0"

__all__ = ["ActivitySampler", "ConvPlan", "describe_plan", "DoseCalculationResult", "DoseCalculator",
           "DosimetryCalculator", "Ga68KernelGenerator", "KernelConvolutionCalculator", "KernelFactory",
           "Lu177KernelGenerator", "Y90KernelGenerator"]
