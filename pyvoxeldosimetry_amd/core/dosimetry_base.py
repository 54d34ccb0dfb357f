"""Plugin contract for dose-calculation methods.

Same contract as ``DosimetryCalculator`` (``pyvoxeldosimetry/core/dosimetry_base.py:8-71``):
constructor ``(radionuclide, tissue_composition, config)``, a ``TypeError`` for
a non-string radionuclide, the two abstract dose methods and ``get_config``.
"""
from __future__ import annotations

from abc import ABC, abstractmethod
from typing import Any, Dict, List, Optional, Tuple


class DosimetryCalculator(ABC):
    """Abstract base class for all dosimetry calculation methods."""

    def __init__(self, radionuclide: str, tissue_composition: Any, config: Optional[Dict[str, Any]] = None):
        self.radionuclide = radionuclide
        self.tissue_composition = tissue_composition
        self.config = config or {}
        self._validate_inputs()

    def _validate_inputs(self):
        if not isinstance(self.radionuclide, str):
            raise TypeError("Radionuclide must be a string")

    @abstractmethod
    def calculate_dose_rate(self, activity_map, voxel_size: Tuple[float, float, float]):
        """Dose-rate map (Gy/s nominal) from an activity map (Bq)."""

    @abstractmethod
    def calculate_absorbed_dose(self, activity_maps: List, time_points: List[float],
                                voxel_size: Tuple[float, float, float]):
        """Absorbed dose (Gy nominal) from a time series of activity maps (hours)."""

    def get_config(self) -> Dict[str, Any]:
        return self.config.copy()
