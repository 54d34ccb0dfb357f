"""File formats around the dose path (NIfTI-1 without nibabel)."""
from .nifti import is_nifti_file  # noqa: F401
