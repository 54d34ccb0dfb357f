"""Dose-point kernels (Y-90, Lu-177, Ga-68) generated on the GPU."""
from .config import load_config, strip_json_comments
from .generators import GENERATORS, Ga68KernelGenerator, Lu177KernelGenerator, Y90KernelGenerator
from .kernel_factory import KernelFactory
from .nuclides import NUCLIDES, SUPPORTED, TISSUES, nuclide_from_config, tissue

__all__ = ["GENERATORS", "Ga68KernelGenerator", "KernelFactory", "Lu177KernelGenerator", "Y90KernelGenerator",
           "NUCLIDES", "SUPPORTED", "TISSUES", "load_config", "nuclide_from_config", "strip_json_comments", "tissue"]
