"""Restated dose-point-kernel generators (FP64 NumPy).

Test infrastructure only (see ``oracle/__init__.py``).

Follows ``pyvoxeldosimetry/data/dose_kernels/y90_kernel.py:20-162``,
``lu177_kernel.py:52-184`` and ``ga68_kernel.py:20-106`` with the constants of
``Y90/Y90.json:13-25``, ``Lu177/Lu177.json:13-27`` and ``Ga68/Ga68.json:13-25``
(the latter two have ``//`` comments, so the reference's ``json.load`` at
``base_kernel.py:34-36`` cannot read them; the values are the ones the files state).

Pinned by (tests/test_oracle.py), independently of these formulas' text:
  * the centre voxel: Y-90 raw is NaN (``inf * 0`` at ``y90_kernel.py:134-138``),
    Y-90 fixed centre == rho*spr*f_tissue (1.0 in water), Lu-177 centre ==
    sum of beta abundances (1.0); Ga-68 raw NaN, fixed centre == f_tissue;
  * far field beyond every beta range equals the published photon point-source
    term ``I exp(-mu r/10)/(4 pi r^2)`` at radii computed from integer offsets
    (Ga-68: the annihilation line plus the 1.077 MeV line, times f_tissue); the
    Ga-68 positron term is zero from 9 E^1.5 f mm on;
  * the grid-centre convention ``g//2`` on even grids (64^3 grid: the NaN sits
    at index 32, and the shipped ``cached_kernels/Y90_water_xy.png`` renders a
    64x64 slice with colour-bar max ~0.90 and a blank (NaN) centre);
  * isotropy: the kernel is invariant under axis permutations and reflections
    about the centre for odd grids.
"""
from __future__ import annotations

import math

import numpy as np

# Y90.json:21-25 / Lu177.json:20-27 -------------------------------------------------
Y90_BETA_MAX = 2.280
Y90_HALF_LIFE_H = 64.1
LU177_BETA_MAX = (0.498, 0.385, 0.177)
LU177_BETA_ABUND = (0.795, 0.089, 0.116)
LU177_GAMMAS = ((0.208, 0.111), (0.113, 0.062))   # (energy MeV, intensity)
LU177_HALF_LIFE_H = 161.52

# y90_kernel.py:59-90 (density, effective_Z, stopping_power_ratio)
Y90_TISSUE = {
    "water": (1.0, 7.42, 1.0),
    "lung": (0.26, 7.41, 1.04),
    "soft_tissue": (1.04, 7.46, 1.04),
    "bone": (1.85, 13.8, 1.15),
    "iodine_contrast": (1.30, 53.0, 1.12),
}
# y90_kernel.py:148-162
Y90_TISSUE_FACTOR = {"water": 1.0, "bone": 1.15, "lung": 1.04,
                     "soft_tissue": 1.04, "iodine_contrast": 1.12}
# lu177_kernel.py:76-117 (density, stopping_power_ratio, mu_by_rho)
LU177_TISSUE = {
    "water": (1.0, 1.0, 0.096),
    "lung": (0.26, 1.04, 0.095),
    "soft_tissue": (1.04, 1.04, 0.097),
    "bone": (1.85, 1.15, 0.110),
    "iodine_contrast": (1.30, 1.12, 0.245),
}


def radius_grid(voxel_size: float, grid) -> np.ndarray:
    """``y90_kernel.py:33-43``: r in mm about the centre index ``g//2``."""
    c = [s // 2 for s in grid]
    x, y, z = np.meshgrid(np.arange(grid[0]) - c[0], np.arange(grid[1]) - c[1],
                          np.arange(grid[2]) - c[2], indexing="ij")
    return np.sqrt((x * voxel_size) ** 2 + (y * voxel_size) ** 2 + (z * voxel_size) ** 2)


def _beta(r, range_water_mm, rho, spr):
    """``y90_kernel.py:93-117`` / ``lu177_kernel.py:129-154``."""
    tissue_range = range_water_mm * (1.0 / rho) * (1.0 / spr)
    out = np.zeros_like(r)
    m = r <= tissue_range
    out[m] = (1 - r[m] / tissue_range) ** 2 * np.exp(-2 * r[m] / tissue_range) * rho * spr
    return out


def y90_kernel(voxel_size: float, grid, tissue: str = "water", fix_centre: bool = False):
    """``Y90KernelGenerator.generate_kernel`` (``y90_kernel.py:20-55``).

    ``fix_centre=False`` reproduces the NaN centre voxel of the reference
    (``:134-138``); ``True`` sets the bremsstrahlung term to 0 at r=0, the
    evident intent of the ``(r > 0)`` factor.
    """
    rho, zeff, spr = Y90_TISSUE.get(tissue, Y90_TISSUE["water"])
    r = radius_grid(voxel_size, grid)
    k = np.zeros(tuple(grid))
    k += _beta(r, 11.0 * Y90_BETA_MAX ** 1.5, rho, spr)
    mu = 0.096 * (rho / 1.0)
    with np.errstate(divide="ignore", invalid="ignore"):
        brems = 0.015 * (zeff / 7.42) ** 2 * rho * np.exp(-mu * r / 10) / (4 * np.pi * r ** 2) * (r > 0)
    if fix_centre:
        brems[r == 0] = 0.0
    k += brems
    k *= Y90_TISSUE_FACTOR.get(tissue, 1.0)
    return k


def lu177_kernel(voxel_size: float, grid, tissue: str = "water"):
    """``Lu177KernelGenerator.generate_kernel`` (``lu177_kernel.py:52-75``)."""
    rho, spr, mu_rho = LU177_TISSUE.get(tissue, LU177_TISSUE["water"])
    r = radius_grid(voxel_size, grid)
    k = np.zeros(tuple(grid))
    for e, a in zip(LU177_BETA_MAX, LU177_BETA_ABUND):
        k += a * _beta(r, 5.0 * e ** 1.5, rho, spr)
    for e, inten in LU177_GAMMAS:
        mu = rho * mu_rho * (0.2 / e) ** 3.2
        g = np.zeros_like(r)
        m = r > 0
        g[m] = inten * np.exp(-mu * r[m] / 10) / (4 * np.pi * (r[m] ** 2))
        k += g
    return k


# Ga68.json:16-25 and ga68_kernel.py:84-106 (tissue scaling factor; attenuation at 511 keV, 1/cm)
GA68_BETA_MAX = 1.899
GA68_ANNIHILATION = 0.511
GA68_GAMMAS = ((1.077, 0.03),)
GA68_HALF_LIFE_H = 1.128
GA68_TISSUE_FACTOR = {"water": 1.0, "lung": 0.3, "soft_tissue": 1.04, "bone": 1.6, "iodine_contrast": 1.3}
GA68_MU = {"water": 0.096, "lung": 0.029, "soft_tissue": 0.099, "bone": 0.172, "iodine_contrast": 0.158}


def ga68_kernel(voxel_size: float, grid, tissue: str = "water", fix_centre: bool = False):
    """``Ga68KernelGenerator.generate_kernel`` (``ga68_kernel.py:20-106``): positron
    beta with range ``9 E^1.5 * f`` mm (``:60-62``, the tissue factor multiplies), the
    511 keV annihilation term and the gamma lines ``I e^{-mu r/10}/(4 pi r^2) (r>0)``
    with ``mu = mu_511(tissue) (0.511/E)^3.2``, the whole kernel times ``f``.
    ``fix_centre=False`` keeps the NaN centre (``inf * 0``); ``True`` gives 0 there."""
    f = GA68_TISSUE_FACTOR.get(tissue, 1.0)
    mu511 = GA68_MU.get(tissue, 0.096)
    r = radius_grid(voxel_size, grid)
    k = np.zeros(tuple(grid))
    rng = 9.0 * GA68_BETA_MAX ** 1.5 * f
    m = r <= rng
    k[m] += (1 - r[m] / rng) ** 2 * np.exp(-2 * r[m] / rng)
    with np.errstate(divide="ignore", invalid="ignore"):
        for e, inten in ((GA68_ANNIHILATION, 1.0),) + GA68_GAMMAS:
            mu = mu511 * (0.511 / e) ** 3.2
            g = inten * np.exp(-mu * r / 10) / (4 * np.pi * r ** 2) * (r > 0)
            if fix_centre:
                g[r == 0] = 0.0
            k += g
    return k * f


def synthetic_kernel(voxel_size: float, grid, length_mm: float = 2.5, core_mm2: float = 0.25):
    """Isotropic ``k(r) ~ exp(-r/2.5 mm)/(r^2 + 0.25)`` normalised to sum 1."""
    r = radius_grid(voxel_size, grid)
    k = np.exp(-r / length_mm) / (r ** 2 + core_mm2)
    return k / k.sum()


def photon_point_source(r_mm: float, intensity: float, mu_per_cm: float) -> float:
    """Published isotropic photon point-source term ``I e^{-mu r/10}/(4 pi r^2)``."""
    return intensity * math.exp(-mu_per_cm * r_mm / 10.0) / (4.0 * math.pi * r_mm * r_mm)
