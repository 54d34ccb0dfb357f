"""Beta spectra, photon lines and range / absorption tables for the Monte Carlo.

Spectrum shapes are the tabulated points of the reference's decay data
(``pyvoxeldosimetry/data/mc_decay/Y90.json:14-17``, ``Lu177.json:14-17``),
interpolated linearly in energy.  The continuous-slowing-down range in water
is the Katz-Penfold relation R(E) = 0.412 E^(1.265 - 0.0954 ln E) g/cm^2
(E in MeV, 0.01-2.5 MeV; Katz & Penfold, Rev. Mod. Phys. 24 (1952) 28).

Photons: the gamma lines per decay are those of the kernel generators
(``Lu177/Lu177.json:23-26``: 0.208 MeV x 0.111, 0.113 MeV x 0.062); Y-90 emits
bremsstrahlung (``Y90/Y90.json:24``), here one photon packet per electron with the
thick-target radiation yield Y(E) = 6e-4 Z E / (1 + 6e-4 Z E) of a medium of atomic
number Z (Evans, The Atomic Nucleus, 1955, ch. 21).  Energy absorption uses the
mass energy-absorption coefficient of liquid water (Hubbell & Seltzer, NIST
Standard Reference Database 126), interpolated log-log.
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


PHOTON_LINES = {
    "Y90": (),
    "Lu177": ((0.208, 0.111), (0.113, 0.062)),
}
BREMS_Z = {"Y90": 7.42, "Lu177": 0.0}      # bremsstrahlung packets (Z of water) or none

# mu_en / rho of liquid water, cm^2/g (Hubbell & Seltzer, NIST SRD 126)
WATER_MU_EN = (
    (0.010, 4.944), (0.015, 1.374), (0.020, 0.5503), (0.030, 0.1557), (0.040, 0.06947), (0.050, 0.04223),
    (0.060, 0.03190), (0.080, 0.02597), (0.100, 0.02546), (0.150, 0.02764), (0.200, 0.02967), (0.300, 0.03192),
    (0.400, 0.03279), (0.500, 0.03299), (0.600, 0.03284), (0.800, 0.03206), (1.000, 0.03103), (1.250, 0.02965),
    (1.500, 0.02833), (2.000, 0.02608), (3.000, 0.02281),
)


def water_mu_en(e_mev):
    """mu_en/rho of water (cm^2/g), log-log interpolation, clamped to the table ends."""
    le = np.log([e for e, _ in WATER_MU_EN])
    lm = np.log([m for _, m in WATER_MU_EN])
    x = np.clip(np.log(np.asarray(e_mev, dtype=np.float64)), le[0], le[-1])
    return np.exp(np.interp(x, le, lm))


def mu_en_table(n: int = 256):
    """(table of mu_en/rho on log-spaced energies, log E_lo, log E_hi) for the device."""
    lo, hi = np.log(WATER_MU_EN[0][0]), np.log(WATER_MU_EN[-1][0])
    return water_mu_en(np.exp(np.linspace(lo, hi, n))), float(lo), float(hi)


def brems_yield(e_mev, z: float):
    """Thick-target bremsstrahlung yield Y(E) = 6e-4 Z E / (1 + 6e-4 Z E)."""
    x = 6.0e-4 * z * np.asarray(e_mev, dtype=np.float64)
    return x / (1.0 + x)
