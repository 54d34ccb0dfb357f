"""Analytic dose-point-kernel generators evaluated on the GPU in FP64.

Mirrors ``Y90KernelGenerator`` (``pyvoxeldosimetry/data/dose_kernels/y90_kernel.py:7-162``),
``Lu177KernelGenerator`` (``lu177_kernel.py:38-190``) and ``Ga68KernelGenerator``
(``ga68_kernel.py:7-106``): the same kernel ``K(r)`` on the same grid (centre
``g//2``), built as a list of terms that ``pvd_dpk_generate`` sums on the device.
Every generator takes an optional ``config_path``: a nuclide JSON file in the
reference's layout (``base_kernel.py:22-36``), ``//`` comments allowed, whose
constants replace this package's own.
"""
from __future__ import annotations

import ctypes
from typing import List, Tuple

import numpy as np

from .. import _native as nat
from .config import load_config
from .nuclides import GA68_MU_511, GA68_TISSUE_FACTOR, NUCLIDES, Y90_TISSUE_FACTOR, nuclide_from_config, tissue


def _beta_term(amp: float, range_water_mm: float, rho: float, spr: float) -> nat.DpkTerm:
    # tissue range = max_range * (1/rho) * (1/spr)  (y90_kernel.py:103-104, lu177_kernel.py:141-142)
    return nat.DpkTerm(nat.PVD_DPK_BETA, 0, amp, range_water_mm * (1.0 / rho) * (1.0 / spr))


class _DeviceGenerator:
    symbol = ""

    def __init__(self, tissue_type: str, config_path=None):
        if not isinstance(tissue_type, str):
            raise TypeError("tissue_type must be a string")
        self.tissue_type = tissue_type
        if config_path is not None:
            self.config = load_config(config_path)
            self.nuclide = nuclide_from_config(self.config)
            if self.nuclide.symbol != self.symbol:
                raise ValueError(f"{config_path} describes {self.nuclide.symbol}, not {self.symbol}")
        else:
            self.config = None
            self.nuclide = NUCLIDES[self.symbol]
        self.tissue_properties = tissue(tissue_type)

    def terms(self) -> Tuple[List[nat.DpkTerm], float]:
        raise NotImplementedError

    def generate_kernel_device(self, voxel_size: float, grid_size, device=None):
        """FP64 kernel as a torch tensor on the GPU."""
        import torch
        nat.require_device()
        lib = nat.load()
        g = tuple(int(s) for s in grid_size)
        if len(g) != 3 or min(g) < 1:
            raise ValueError(f"grid_size must be 3 positive ints, got {grid_size}")
        dev = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
        out = torch.empty(g, dtype=torch.float64, device=dev)
        terms, factor = self.terms()
        arr = (nat.DpkTerm * len(terms))(*terms)
        with torch.cuda.device(dev):
            nat.check(lib.pvd_dpk_generate(float(voxel_size), g[0], g[1], g[2], arr, len(terms), float(factor),
                                           ctypes.c_void_p(out.data_ptr()),
                                           ctypes.c_void_p(nat.current_stream_handle(dev))),
                      "pvd_dpk_generate")
        return out

    def generate_kernel(self, voxel_size: float, grid_size) -> np.ndarray:
        """Same contract as the reference: an FP64 NumPy array."""
        return self.generate_kernel_device(voxel_size, grid_size).cpu().numpy()


class Y90KernelGenerator(_DeviceGenerator):
    """``y90_kernel.py:20-55``: beta (Cole-type) + bremsstrahlung, times a tissue factor.

    ``fix_centre=False`` keeps the reference's NaN centre voxel (``:134-138``:
    ``.../(4 pi r^2) * (r > 0)`` is ``inf * 0`` at r = 0); ``True`` gives 0 there.
    """
    symbol = "Y90"

    def __init__(self, tissue_type: str, fix_centre: bool = False, config_path=None):
        super().__init__(tissue_type, config_path)
        self.fix_centre = bool(fix_centre)

    def terms(self):
        t = self.tissue_properties
        n = self.nuclide
        rho, spr = t.density, t.stopping_power_ratio
        terms = [_beta_term(rho * spr, n.beta_range_coeff * n.beta_max_mev[0] ** 1.5, rho, spr)]
        if n.bremsstrahlung:
            relative_yield = (t.effective_z / 7.42) ** 2                 # :128
            mu = 0.096 * (rho / 1.0)                                       # :142-146
            terms.append(nat.DpkTerm(nat.PVD_DPK_PHOTON, 0 if self.fix_centre else nat.PVD_DPK_NAN_CENTRE,
                                     0.015 * relative_yield * rho, mu))
        return terms, Y90_TISSUE_FACTOR.get(self.tissue_type, 1.0)


class Lu177KernelGenerator(_DeviceGenerator):
    """``lu177_kernel.py:52-75``: three betas weighted by abundance + two photon lines."""
    symbol = "Lu177"

    def terms(self):
        t = self.tissue_properties
        n = self.nuclide
        rho, spr = t.density, t.stopping_power_ratio
        terms = [_beta_term(a * rho * spr, n.beta_range_coeff * e ** 1.5, rho, spr)
                 for e, a in zip(n.beta_max_mev, n.beta_abundance)]
        for e, inten in n.gammas:                                          # :156-184
            mu = rho * t.mu_by_rho * (0.2 / e) ** 3.2
            terms.append(nat.DpkTerm(nat.PVD_DPK_PHOTON, 0, inten, mu))
        return terms, 1.0


class Ga68KernelGenerator(_DeviceGenerator):
    """``ga68_kernel.py:20-106``: positron beta + 511 keV annihilation photons + gamma lines,
    times the tissue scaling factor.

    The positron range is 9 E^1.5 mm times that factor (``:60-62``: multiplied, as the
    reference does); the photon terms use the tissue's 511 keV coefficient scaled by
    (0.511/E)^3.2 (``:96-106``).  ``fix_centre=False`` keeps the reference's NaN centre
    (``exp(0)/(4 pi 0) * (r > 0)`` is ``inf * 0``); ``True`` gives 0 there.
    """
    symbol = "Ga68"

    def __init__(self, tissue_type: str, fix_centre: bool = False, config_path=None):
        super().__init__(tissue_type, config_path)
        self.fix_centre = bool(fix_centre)

    def terms(self):
        n = self.nuclide
        f = GA68_TISSUE_FACTOR.get(self.tissue_type, 1.0)
        mu511 = GA68_MU_511.get(self.tissue_type, 0.096)
        centre = 0 if self.fix_centre else nat.PVD_DPK_NAN_CENTRE
        terms = [nat.DpkTerm(nat.PVD_DPK_BETA, 0, 1.0, n.beta_range_coeff * n.beta_max_mev[0] ** 1.5 * f)]
        e_ann = n.annihilation_mev or 0.511
        terms.append(nat.DpkTerm(nat.PVD_DPK_PHOTON, centre, 1.0, mu511 * (0.511 / e_ann) ** 3.2))
        for e, inten in n.gammas:
            terms.append(nat.DpkTerm(nat.PVD_DPK_PHOTON, centre, inten, mu511 * (0.511 / e) ** 3.2))
        return terms, f


GENERATORS = {"Y90": Y90KernelGenerator, "Lu177": Lu177KernelGenerator, "Ga68": Ga68KernelGenerator}
