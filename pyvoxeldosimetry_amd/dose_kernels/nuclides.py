"""Nuclide and tissue constants of the analytic dose-point kernels.

Nuclide constants come from nuclide JSON files in the reference's layout
(``pyvoxeldosimetry/data/dose_kernels/<Nuclide>/<Nuclide>.json``, read by
``base_kernel.py:34-36``); this package ships its own copies of the values in
``dose_kernels/data`` and ``nuclide_from_config`` builds a ``Nuclide`` from any
such file (``//`` comments allowed, see ``config.py``):
  * Y-90 (``Y90/Y90.json:13-25``): half-life 64.1 h, beta max 2.280 MeV,
    bremsstrahlung on, scaling_factor 1e-12;
  * Lu-177 (``Lu177/Lu177.json:13-33``): half-life 161.52 h, three betas, two
    gamma lines;
  * Ga-68 (``Ga68/Ga68.json:13-30``): half-life 1.128 h, beta+ max 1.899 MeV,
    511 keV annihilation photons, one 1.077 MeV gamma line.
The tissue tables are the generators' own (``y90_kernel.py:57-91``,
``lu177_kernel.py:76-117``, ``ga68_kernel.py:84-106``), kept identical so kernels match.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Dict, List, Tuple

from .config import packaged_config


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

# Ga-68 tissue scaling of the positron range and of the whole kernel, and linear
# attenuation at 511 keV in 1/cm; anything else 1.0 / 0.096 (ga68_kernel.py:84-106)
GA68_TISSUE_FACTOR: Dict[str, float] = {"water": 1.0, "lung": 0.3, "soft_tissue": 1.04, "bone": 1.6,
                                        "iodine_contrast": 1.3}
GA68_MU_511: Dict[str, float] = {"water": 0.096, "lung": 0.029, "soft_tissue": 0.099, "bone": 0.172,
                                 "iodine_contrast": 0.158}

# CSDA range coefficient of each generator's beta model, range = coeff * E^1.5 mm in
# water (y90_kernel.py:101, lu177_kernel.py:139, ga68_kernel.py:62): code, not file data
BETA_RANGE_COEFF: Dict[str, float] = {"Y90": 11.0, "Lu177": 5.0, "Ga68": 9.0}


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
    positron: bool = False
    annihilation_mev: float = 0.0
    extra: Dict[str, float] = field(default_factory=dict)


def _as_tuple(v) -> Tuple[float, ...]:
    return tuple(float(x) for x in (v if isinstance(v, (list, tuple)) else [v]))


def nuclide_from_config(cfg: dict) -> Nuclide:
    """A ``Nuclide`` from a parsed nuclide JSON of the reference's layout."""
    nuc = cfg["nuclide"]
    pe = nuc.get("particle_energies", {})
    symbol = nuc["symbol"]
    positron = "beta+_max" in pe
    beta_max = _as_tuple(pe["beta+_max"] if positron else pe["beta-_max"])
    abund = _as_tuple(pe.get("beta-_abundance", [1.0] * len(beta_max)))
    if len(abund) != len(beta_max):
        raise ValueError(f"{symbol}: {len(beta_max)} beta energies but {len(abund)} abundances")
    gammas = tuple((float(g["energy"]), float(g["intensity"])) for g in pe.get("gamma_lines", []))
    kern = cfg.get("kernel", {})
    return Nuclide(name=nuc["name"], symbol=symbol, half_life_h=float(nuc["half_life"]), beta_max_mev=beta_max,
                   beta_abundance=abund, beta_range_coeff=BETA_RANGE_COEFF.get(symbol, 0.0), gammas=gammas,
                   bremsstrahlung=bool(pe.get("bremsstrahlung", False)),
                   scaling_factor=float(kern.get("scaling_factor", 1.0e-12)),
                   default_grid=tuple(int(g) for g in kern.get("grid_size", (101, 101, 101))),
                   positron=positron, annihilation_mev=float(pe.get("annihilation", 0.0)))


NUCLIDES: Dict[str, Nuclide] = {s: nuclide_from_config(packaged_config(s)) for s in ("Y90", "Lu177", "Ga68")}

SUPPORTED: List[str] = list(NUCLIDES)


def tissue(name: str) -> Tissue:
    """Unknown tissue names fall back to water (y90_kernel.py:91, lu177_kernel.py:117)."""
    return TISSUES.get(name, TISSUES["water"])
