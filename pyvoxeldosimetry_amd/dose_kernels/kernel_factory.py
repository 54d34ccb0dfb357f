"""Kernel registry with a correctly keyed cache.

Mirrors ``KernelFactory`` (``pyvoxeldosimetry/data/dose_kernels/kernel_factory.py:12-74``):
same ``get_kernel`` signature, same ``ValueError`` for unknown nuclides, same
per-nuclide default grids.  Differences: the cache is in memory and keyed by
every input that changes the kernel (the reference keys on nuclide+tissue
only, ``:66-69``, and writes into its package tree), and kernels are computed
on the GPU.
"""
from __future__ import annotations

from typing import Dict, Optional, Tuple

import numpy as np

from .generators import GENERATORS
from .nuclides import NUCLIDES


class KernelFactory:
    _default_grid_sizes: Dict[str, Tuple[int, int, int]] = {k: v.default_grid for k, v in NUCLIDES.items()}

    def __init__(self):
        self._cache: Dict[tuple, object] = {}

    @staticmethod
    def supported() -> list:
        return list(GENERATORS)

    def _check(self, nuclide: str):
        if nuclide not in GENERATORS:
            raise ValueError(f"Unsupported nuclide: {nuclide}. Supported: {list(GENERATORS)}")

    def get_kernel_device(self, nuclide: str, tissue_type: str, voxel_size: float = 1.0,
                          grid_size: Optional[Tuple[int, int, int]] = None, force_regenerate: bool = False,
                          fix_centre: bool = False, device=None):
        """FP64 kernel as a torch tensor on the GPU (cached)."""
        self._check(nuclide)
        if grid_size is None:
            grid_size = self._default_grid_sizes.get(nuclide, (101, 101, 101))
        grid = tuple(int(s) for s in grid_size)
        key = (nuclide, tissue_type, float(voxel_size), grid, bool(fix_centre), str(device))
        if not force_regenerate and key in self._cache:
            return self._cache[key]
        cls = GENERATORS[nuclide]
        gen = cls(tissue_type, fix_centre=fix_centre) if nuclide in ("Y90", "Ga68") else cls(tissue_type)
        k = gen.generate_kernel_device(voxel_size, grid, device=device)
        self._cache[key] = k
        return k

    def get_kernel(self, nuclide: str, tissue_type: str, voxel_size: float = 1.0,
                   grid_size: Optional[Tuple[int, int, int]] = None, force_regenerate: bool = False,
                   fix_centre: bool = False) -> np.ndarray:
        """Reference signature (``kernel_factory.py:37-74``); returns FP64 NumPy."""
        return self.get_kernel_device(nuclide, tissue_type, voxel_size, grid_size, force_regenerate,
                                      fix_centre).cpu().numpy()
