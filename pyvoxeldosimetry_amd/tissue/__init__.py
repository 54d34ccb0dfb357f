"""HU -> tissue classes and mass density on the GPU (the producer of the
density map that the dose epilogue consumes)."""
from .composition import (HU_CLASSES, HU_MATERIALS, TissueComposition, TissueManager, TissueProperties,
                          density_from_hu)

__all__ = ["HU_CLASSES", "HU_MATERIALS", "TissueComposition", "TissueManager", "TissueProperties",
           "density_from_hu"]
