"""Beta spectra and electron range tables for the Monte Carlo.

Spectrum shapes are the tabulated points of the reference's decay data
(``pyvoxeldosimetry/data/mc_decay/Y90.json:14-17``, ``Lu177.json:14-17``),
interpolated linearly in energy.  The continuous-slowing-down range in water
is the Katz-Penfold relation R(E) = 0.412 E^(1.265 - 0.0954 ln E) g/cm^2
(E in MeV, 0.01-2.5 MeV; Katz & Penfold, Rev. Mod. Phys. 24 (1952) 28).
"""
from __future__ import annotations

import numpy as np

BETA_SPECTRA = {
    "Y90": ([0.0, 0.2, 0.4, 0.6, 0.8, 1.0, 1.2, 1.4, 1.6, 1.8, 2.0, 2.2, 2.280],
            [0.0, 0.12, 0.22, 0.30, 0.36, 0.40, 0.42, 0.38, 0.32, 0.24, 0.15, 0.06, 0.0]),
    "Lu177": ([0.0, 0.1, 0.2, 0.3, 0.4, 0.498],
              [0.0, 0.35, 0.48, 0.42, 0.22, 0.0]),
}

KATZ_A, KATZ_B, KATZ_C = 0.412, 1.265, 0.0954


def katz_range_mm(e_mev):
    """Water CSDA range in mm."""
    e = np.maximum(np.asarray(e_mev, dtype=np.float64), 1e-4)
    return 10.0 * KATZ_A * e ** (KATZ_B - KATZ_C * np.log(e))


def inverse_cdf(nuclide: str, n: int = 1025) -> np.ndarray:
    """Energies E(u) at u = i/(n-1) for the piecewise-linear spectrum."""
    e_pts, p_pts = (np.asarray(v, dtype=np.float64) for v in BETA_SPECTRA[nuclide])
    fine = np.linspace(e_pts[0], e_pts[-1], 20001)
    pdf = np.interp(fine, e_pts, p_pts)
    cdf = np.concatenate([[0.0], np.cumsum(0.5 * (pdf[1:] + pdf[:-1]) * np.diff(fine))])
    cdf /= cdf[-1]
    u = np.linspace(0.0, 1.0, n)
    return np.interp(u, cdf, fine)


def mean_energy(nuclide: str) -> float:
    e_pts, p_pts = (np.asarray(v, dtype=np.float64) for v in BETA_SPECTRA[nuclide])
    fine = np.linspace(e_pts[0], e_pts[-1], 20001)
    pdf = np.interp(fine, e_pts, p_pts)
    return float(np.trapezoid(fine * pdf, fine) / np.trapezoid(pdf, fine))


def energy_of_range(e_max: float, n: int = 4096):
    """(table E(r) on r in [0, r_max], r_max) with r the water range in mm."""
    e_grid = np.geomspace(1e-4, e_max, 200001)
    r_grid = katz_range_mm(e_grid)
    if np.any(np.diff(r_grid) <= 0):
        raise ValueError("range-energy relation is not monotonic on the table interval")
    r_max = float(r_grid[-1])
    r = np.linspace(0.0, r_max, n)
    e = np.interp(r, np.concatenate([[0.0], r_grid]), np.concatenate([[0.0], e_grid]))
    return e, r_max
