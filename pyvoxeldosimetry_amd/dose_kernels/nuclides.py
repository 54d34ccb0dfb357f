"""Nuclide and tissue constants of the analytic dose-point kernels.

Values as stated by the reference's data files (tables of physical constants,
kept identical so kernels match):
  * ``pyvoxeldosimetry/data/dose_kernels/Y90/Y90.json:13-25`` (half-life 64.1 h,
    beta max 2.280 MeV, bremsstrahlung on, scaling_factor 1e-12 at ``:31``);
  * ``Lu177/Lu177.json:13-33`` (half-life 161.52 h, betas, gammas; the file
    carries ``//`` comments, so these are the values it states);
  * tissue tables ``y90_kernel.py:57-91`` and ``lu177_kernel.py:76-117``, tissue
    factors ``y90_kernel.py:148-162``.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Dict, List, Tuple


@dataclass(frozen=True)
class Tissue:
    density: float
    effective_z: float
    stopping_power_ratio: float
    mu_by_rho: float          # cm^2/g at 0.2 MeV (Lu-177 photon model)


TISSUES: Dict[str, Tissue] = {
    "water": Tissue(1.0, 7.42, 1.0, 0.096),
    "lung": Tissue(0.26, 7.41, 1.04, 0.095),
    "soft_tissue": Tissue(1.04, 7.46, 1.04, 0.097),
    "bone": Tissue(1.85, 13.8, 1.15, 0.110),
    "iodine_contrast": Tissue(1.30, 53.0, 1.12, 0.245),
}

# Y-90 kernel-level factor per tissue name; anything else is 1.0 (y90_kernel.py:148-162)
Y90_TISSUE_FACTOR: Dict[str, float] = {"water": 1.0, "bone": 1.15, "lung": 1.04,
                                       "soft_tissue": 1.04, "iodine_contrast": 1.12}


@dataclass(frozen=True)
class Nuclide:
    name: str
    symbol: str
    half_life_h: float
    beta_max_mev: Tuple[float, ...]
    beta_abundance: Tuple[float, ...]
    beta_range_coeff: float            # CSDA range = coeff * E^1.5 mm in water
    gammas: Tuple[Tuple[float, float], ...] = ()   # (energy MeV, intensity)
    bremsstrahlung: bool = False
    scaling_factor: float = 1.0e-12
    default_grid: Tuple[int, int, int] = (101, 101, 101)
    extra: Dict[str, float] = field(default_factory=dict)


NUCLIDES: Dict[str, Nuclide] = {
    "Y90": Nuclide("Yttrium-90", "Y90", 64.1, (2.280,), (1.0,), 11.0, (), True, 1.0e-12, (201, 201, 201)),
    "Lu177": Nuclide("Lutetium-177", "Lu177", 161.52, (0.498, 0.385, 0.177), (0.795, 0.089, 0.116), 5.0,
                     ((0.208, 0.111), (0.113, 0.062)), False, 1.0e-12, (81, 81, 81)),
}

SUPPORTED: List[str] = list(NUCLIDES)


def tissue(name: str) -> Tissue:
    """Unknown tissue names fall back to water (y90_kernel.py:91, lu177_kernel.py:117)."""
    return TISSUES.get(name, TISSUES["water"])
