"""Dose-point kernels (Y-90, Lu-177) generated on the GPU."""
from .generators import GENERATORS, Lu177KernelGenerator, Y90KernelGenerator
from .kernel_factory import KernelFactory
from .nuclides import NUCLIDES, SUPPORTED, TISSUES, tissue

__all__ = ["GENERATORS", "KernelFactory", "Lu177KernelGenerator", "Y90KernelGenerator", "NUCLIDES",
           "SUPPORTED", "TISSUES", "tissue"]
