"""Tissue composition and density from CT on the GPU.

Drop-in for ``pyvoxeldosimetry.tissue.composition`` (``TissueProperties``,
``TissueManager``, ``TissueComposition``, composition.py:11-94) over the HIP
kernel of ``include/pvd_tissue.h``, plus ``density_from_hu``: HU -> material ->
g/cm3 through the ranges of ``core/gate/data/HU_to_material.txt`` and the
densities of ``core/gate/data/GateMaterials.db``.  Its output feeds the
``tissue_densities`` argument of ``DoseCalculator.calculate_dose``.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass
from typing import Any, Dict, Optional

import numpy as np

from .. import _native as nat
from ..arrays import to_device_f32

# composition.py:40-46: inclusive HU ranges, in the reference's dict order
HU_CLASSES = {
    "air": (-1000.0, -900.0),
    "lung": (-900.0, -500.0),
    "soft_tissue": (-100.0, 100.0),
    "bone": (300.0, 3000.0),
    "water": (-10.0, 10.0),
}

# HU_to_material.txt:5-19 (lower bound of each range, material) with the
# GateMaterials.db densities (g/cm3).  The integer ranges of the file become
# [lo_i, lo_{i+1}) so that non-integer HU fall somewhere; HU below -1050 is
# Air, above 3000 is Tooth.  Tissue, Bone and Tooth have no GateMaterials.db
# entry: Tissue takes the reference's soft-tissue density (1.04,
# data/dose_kernels/y90_kernel.py tissue table), Bone its bone density (1.85,
# same table), Tooth 2.75 (declared here).
HU_MATERIALS = (
    (-1050.0, "Air", 0.00129),
    (-949.0, "Lung", 0.26),
    (-119.0, "Adipose", 0.92),
    (-82.0, "Breast", 1.02),
    (-52.0, "Adipose", 0.92),
    (-22.0, "Breast", 1.02),
    (8.0, "Adipose", 0.92),
    (38.0, "Breast", 1.02),
    (60.0, "Tissue", 1.04),
    (80.0, "Muscle", 1.05),
    (121.0, "Liver", 1.06),
    (201.0, "Blood", 1.06),
    (301.0, "SpineBone", 1.42),
    (401.0, "RibBone", 1.92),
    (1001.0, "Bone", 1.85),
    (2001.0, "Tooth", 2.75),
)

METAL_HU = 2000.0        # composition.py:83
FILL_SIGMA = 1.0         # composition.py:91


@dataclass
class TissueProperties:
    """composition.py:11-18."""
    density: float
    electron_density: float
    effective_Z: float
    chemical_composition: Dict[str, float]
    hu_value: float
    description: str


class TissueManager:
    """composition.py:20-35: a name -> TissueProperties table (water only, as in the reference)."""

    def __init__(self):
        self.tissue_db = {
            "water": TissueProperties(density=1.0, electron_density=3.34e23, effective_Z=7.42,
                                      chemical_composition={"H": 0.111894, "O": 0.888106}, hu_value=0,
                                      description="Water reference material"),
        }

    def load_tissue(self, name: str) -> Optional[TissueProperties]:
        return self.tissue_db.get(name)


def _params(handle: bool, metal: float, sigma: float, truncate: float, classes, materials):
    p = nat.TissueParams()
    if len(classes) > nat.PVD_TISSUE_MAX_CLASSES:
        raise ValueError(f"at most {nat.PVD_TISSUE_MAX_CLASSES} tissue classes")
    p.n_classes = len(classes)
    for i, (lo, hi) in enumerate(classes):
        p.class_lo[i] = lo
        p.class_hi[i] = hi
    p.n_materials = len(materials)
    for i, (lo, _, rho) in enumerate(materials):
        p.mat_lo[i] = lo
        p.mat_density[i] = rho
    p.handle_artifacts = 1 if handle else 0
    p.metal_hu = metal
    p.sigma = sigma
    p.truncate = truncate
    return p


class TissueComposition:
    """Tissue class maps of a CT image (composition.py:37-94), computed on the GPU.

    Config keys (all optional): ``metal_threshold`` (HU, default 2000),
    ``sigma`` (voxels, default 1), ``truncate`` (default 4, scipy's),
    ``device``.
    """

    def __init__(self, config: Dict[str, Any] = None):
        self.config = config or {}
        self.tissue_types = {k: {"hu_range": v} for k, v in HU_CLASSES.items()}

    # ------------------------------------------------------------------ GPU pass
    def map(self, ct_image, handle_artifacts: bool = True, corrected: bool = True, classes: bool = True,
            density: bool = True):
        """One pass over the CT: (corrected HU, class bits uint8, density g/cm3) as
        GPU tensors (None for outputs not asked for).  Bit c of the class bits is
        the c-th entry of ``tissue_types``."""
        import torch
        nat.require_device()
        lib = nat.load()
        dev = self.config.get("device")
        ct, _ = to_device_f32(ct_image, torch.device(dev) if dev is not None else None)
        if ct.dim() != 3:
            raise ValueError(f"CT image must be 3-D, got shape {tuple(ct.shape)}")
        ranges = [tuple(v["hu_range"]) for v in self.tissue_types.values()]
        p = _params(handle_artifacts, float(self.config.get("metal_threshold", METAL_HU)),
                    float(self.config.get("sigma", FILL_SIGMA)), float(self.config.get("truncate", 4.0)),
                    ranges, HU_MATERIALS)
        out_c = torch.empty_like(ct) if corrected else None
        out_b = torch.empty(ct.shape, dtype=torch.uint8, device=ct.device) if classes else None
        out_d = torch.empty_like(ct) if density else None
        ptr = lambda t: ctypes.c_void_p(t.data_ptr() if t is not None else 0)  # noqa: E731
        with torch.cuda.device(ct.device):
            nat.check(lib.pvd_tissue_map(ctypes.byref(p), ptr(ct), *ct.shape, ptr(out_c), ptr(out_b), ptr(out_d),
                                         ctypes.c_void_p(nat.current_stream_handle(ct.device))), "pvd_tissue_map")
        return out_c, out_b, out_d

    # ------------------------------------------------------------------ reference API
    def calculate_composition(self, ct_image, handle_artifacts: bool = True):
        """Dict tissue name -> 0/1 map (composition.py:48-70).  NumPy in: float64
        NumPy maps, as the reference; GPU tensor in: float32 GPU tensors."""
        import torch
        _, bits, _ = self.map(ct_image, handle_artifacts, corrected=False, classes=True, density=False)
        out = {}
        for c, name in enumerate(self.tissue_types):
            m = ((bits >> c) & 1).to(torch.float32)
            out[name] = m if isinstance(ct_image, torch.Tensor) and ct_image.is_cuda else \
                m.cpu().numpy().astype(np.float64)
        return out

    def _handle_artifacts(self, ct_image):
        """The metal-filled image (composition.py:72-94)."""
        import torch
        corr, _, _ = self.map(ct_image, True, corrected=True, classes=False, density=False)
        if isinstance(ct_image, torch.Tensor) and ct_image.is_cuda:
            return corr
        return corr.cpu().numpy().astype(np.asarray(ct_image).dtype if np.asarray(ct_image).dtype.kind == "f"
                                         else np.float64)


def density_from_hu(ct_image, handle_artifacts: bool = True, config: Dict[str, Any] = None):
    """Mass density (g/cm3, float32) of every voxel of a CT in HU, after the
    metal fill.  NumPy in -> NumPy out; GPU tensor in -> GPU tensor out."""
    import torch
    _, _, d = TissueComposition(config).map(ct_image, handle_artifacts, corrected=False, classes=False,
                                            density=True)
    if isinstance(ct_image, torch.Tensor) and ct_image.is_cuda:
        return d
    return d.cpu().numpy()
