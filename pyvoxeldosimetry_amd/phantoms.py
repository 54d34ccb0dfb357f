"""Seeded synthetic workloads (activity, density, timepoints) for tests and benchmarks.

``create_spherical_activity`` is the workload generator of the reference's
examples (``examples/kernel_convolution_example.py:10-18``: inclusive sphere
``(i-c)^2+(j-c)^2+(k-c)^2 <= r^2``) without the triple Python loop; the
time series follows ``examples/time_integrated_dose.py:26-35``.  The phantom
builders produce the benchmark configurations:

  small      64^3 at 4.42 mm, sphere r=10 at 1 MBq/mL + 5% Poisson noise, water
  ct         256x256x200 at 2.0 mm, liver ellipsoid 0.1 MBq/mL + 3 lesions
             1-5 MBq/mL, synthetic CT in HU (soft tissue 61..69, lungs around
             -700, spine 1200..1800) whose density map is the HU -> material ->
             g/cc table of the tissue producer (Tissue 1.04, Lung 0.26, Bone 1.85)
  batch      48 phantoms, sizes U{96..384}^2 x U{96..320}, 4 timepoints with
             mono-exponential washout (T_eff 20-150 h), Y-90 or Lu-177
  torso      512x512x400 at 0.98 mm, torso with lungs/spine and 10 lesions,
             6 timepoints

Activities are in Bq per voxel (concentration x voxel volume); densities in g/cc.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import List, Sequence, Tuple

import numpy as np


def create_spherical_activity(size=(64, 64, 64), center=None, radius=10, activity=1000.0,
                              dtype=np.float64) -> np.ndarray:
    if center is None:
        center = [s // 2 for s in size]
    x, y, z = np.ogrid[:size[0], :size[1], :size[2]]
    mask = (x - center[0]) ** 2 + (y - center[1]) ** 2 + (z - center[2]) ** 2 <= radius ** 2
    out = np.zeros(size, dtype=dtype)
    out[mask] = activity
    return out


def create_time_series(initial_map, half_life, time_points) -> List[np.ndarray]:
    lam = np.log(2) / half_life
    return [initial_map * np.exp(-lam * t) for t in time_points]


def bq_per_voxel(mbq_per_ml: float, voxel_mm: float) -> float:
    return mbq_per_ml * 1.0e6 * (voxel_mm ** 3) / 1000.0


def _ellipsoid(shape, center, semi, dtype=bool):
    x, y, z = np.ogrid[:shape[0], :shape[1], :shape[2]]
    return (((x - center[0]) / semi[0]) ** 2 + ((y - center[1]) / semi[1]) ** 2
            + ((z - center[2]) / semi[2]) ** 2) <= 1.0


def _fill_sphere(arr, center, radius, value):
    lo = [max(0, int(math.floor(c - radius))) for c in center]
    hi = [min(s, int(math.ceil(c + radius)) + 1) for c, s in zip(center, arr.shape)]
    x, y, z = np.ogrid[lo[0]:hi[0], lo[1]:hi[1], lo[2]:hi[2]]
    m = (x - center[0]) ** 2 + (y - center[1]) ** 2 + (z - center[2]) ** 2 <= radius ** 2
    sub = arr[lo[0]:hi[0], lo[1]:hi[1], lo[2]:hi[2]]
    sub[m] = value


@dataclass
class Phantom:
    name: str
    voxel_mm: float
    activity: List[np.ndarray]            # one map per timepoint (float32)
    density: np.ndarray | None            # float32 or None
    time_points_h: List[float]
    nuclide: str
    kernel_grid: Tuple[int, int, int]
    meta: dict = field(default_factory=dict)

    @property
    def shape(self):
        return self.activity[0].shape


def small_phantom(seed: int = 1) -> Phantom:
    """64^3 at 4.42 mm: sphere r=10 at 1 MBq/mL with 5% Poisson noise, water."""
    rng = np.random.default_rng(seed)
    v = 4.42
    a0 = bq_per_voxel(1.0, v)
    sphere = create_spherical_activity((64, 64, 64), (32, 32, 32), 10, 1.0) > 0
    act = np.zeros((64, 64, 64))
    act[sphere] = a0 * rng.poisson(400.0, size=int(sphere.sum())) / 400.0
    return Phantom("small", v, [act.astype(np.float32)], np.ones((64, 64, 64), np.float32), [2.0], "Y90",
                   (11, 11, 11), {"decay_h": 2.0})


def _body_density(shape, lungs, spine, soft=1.0, lung_rho=0.3, bone_rho=1.85, body=None):
    rho = np.full(shape, soft, dtype=np.float32)
    if body is not None:
        rho[~body] = soft
    for c, s in lungs:
        rho[_ellipsoid(shape, c, s)] = lung_rho
    (cx, cy), r = spine
    x, y = np.ogrid[:shape[0], :shape[1]]
    m2 = (x - cx) ** 2 + (y - cy) ** 2 <= r ** 2
    rho[m2, :] = bone_rho
    return rho


def hu_to_density(hu) -> np.ndarray:
    """g/cc of a CT in HU by the tissue producer's HU -> material table (material i
    covers [lo_i, lo_{i+1}), HU_to_material.txt with GateMaterials.db densities): the
    values pvd_tissue_map writes, computed on the host."""
    from .tissue.composition import HU_MATERIALS
    lo = np.array([m[0] for m in HU_MATERIALS], dtype=np.float64)
    rho = np.array([m[2] for m in HU_MATERIALS], dtype=np.float32)
    idx = np.clip(np.searchsorted(lo, np.asarray(hu, np.float64), side="right") - 1, 0, len(lo) - 1)
    return rho[idx]


def ct_phantom(seed: int = 2) -> Phantom:
    """256x256x200 at 2 mm: liver 0.1 MBq/mL + 3 lesions 1-5 MBq/mL; a synthetic CT in
    HU (meta["hu"]) with lungs and a spine, and its density map (Tissue 1.04, Lung
    0.26, Bone 1.85 g/cc by the HU -> material table)."""
    rng = np.random.default_rng(seed)
    shape = (256, 256, 200)
    v = 2.0
    act = np.zeros(shape, dtype=np.float32)
    liver_c, liver_s = (128, 128, 100), (80, 60, 50)
    act[_ellipsoid(shape, liver_c, liver_s)] = bq_per_voxel(0.1, v)
    for _ in range(3):
        r = int(rng.integers(5, 21))
        # centre inside the liver, sphere kept inside the grid
        while True:
            u = rng.uniform(-0.6, 0.6, size=3)
            c = [liver_c[i] + u[i] * liver_s[i] for i in range(3)]
            if all(r <= c[i] <= shape[i] - 1 - r for i in range(3)):
                break
        _fill_sphere(act, c, r, bq_per_voxel(float(rng.uniform(1.0, 5.0)), v))
    lungs = [((128, 70, 160), (70, 40, 35)), ((128, 186, 160), (70, 40, 35))]
    # HU kept inside each material's range of the table, so the density map stays
    # piecewise constant: soft tissue [60, 80) -> Tissue, lungs [-949, -119) -> Lung,
    # spine [1001, 2001) -> Bone
    hu = rng.uniform(61.0, 69.0, size=shape).astype(np.float32)
    for c, sz in lungs:
        m = _ellipsoid(shape, c, sz)
        hu[m] = np.clip(rng.normal(-700.0, 40.0, size=int(m.sum())), -900.0, -500.0)
    x, y = np.ogrid[:shape[0], :shape[1]]
    spine = (x - 200) ** 2 + (y - 128) ** 2 <= 12 ** 2
    hu[spine, :] = rng.uniform(1200.0, 1800.0, size=(int(spine.sum()), shape[2]))
    rho = hu_to_density(hu)
    return Phantom("ct", v, [act], rho, [0.0], "Lu177", (31, 31, 31), {"hu": hu})


@dataclass
class BatchSpec:
    shape: Tuple[int, int, int]
    t_eff_h: float
    nuclide: str
    kernel_grid: Tuple[int, int, int]
    seed: int


BATCH_TIMEPOINTS_H = [4.0, 24.0, 72.0, 144.0]


def batch_specs(seed: int = 3, count: int = 48) -> List[BatchSpec]:
    """48 uneven phantoms: nx, ny ~ U{96..384}, nz ~ U{96..320}; Y-90 (11^3) or Lu-177 (31^3)."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        shape = (int(rng.integers(96, 385)), int(rng.integers(96, 385)), int(rng.integers(96, 321)))
        t_eff = float(rng.uniform(20.0, 150.0))
        y90 = bool(rng.integers(0, 2))
        out.append(BatchSpec(shape, t_eff, "Y90" if y90 else "Lu177", (11, 11, 11) if y90 else (31, 31, 31),
                             seed * 1000 + i))
    return out


def batch_phantom(spec: BatchSpec, voxel_mm: float = 2.0) -> Phantom:
    """One batch member: organ ellipsoid + 1-3 lesions, activity decaying with T_eff."""
    rng = np.random.default_rng(spec.seed)
    shape = spec.shape
    act0 = np.zeros(shape, dtype=np.float32)
    c = [s / 2 for s in shape]
    semi = [0.3 * s for s in shape]
    act0[_ellipsoid(shape, c, semi)] = bq_per_voxel(0.1, voxel_mm)
    for _ in range(int(rng.integers(1, 4))):
        r = float(min(int(rng.integers(5, 21)), 0.2 * min(shape)))
        u = rng.uniform(-0.5, 0.5, size=3)
        _fill_sphere(act0, [c[i] + u[i] * semi[i] for i in range(3)], r, bq_per_voxel(float(rng.uniform(1, 5)), voxel_mm))
    lam = math.log(2.0) / spec.t_eff_h
    acts = [(act0 * np.float32(math.exp(-lam * t))).astype(np.float32) for t in BATCH_TIMEPOINTS_H]
    return Phantom("batch", voxel_mm, acts, None, list(BATCH_TIMEPOINTS_H), spec.nuclide, spec.kernel_grid,
                   {"t_eff_h": spec.t_eff_h})


TORSO_TIMEPOINTS_H = [1.0, 4.0, 24.0, 48.0, 96.0, 168.0]


def torso_phantom(seed: int = 5, shape=(512, 512, 400), voxel_mm: float = 0.98) -> Phantom:
    """Whole torso: soft tissue 1.04, lungs 0.26, spine 1.85, 10 lesions, 6 timepoints."""
    rng = np.random.default_rng(seed)
    sx, sy, sz = shape
    body = _ellipsoid(shape, (sx / 2, sy / 2, sz / 2), (0.45 * sx, 0.35 * sy, 0.6 * sz))
    act0 = np.zeros(shape, dtype=np.float32)
    act0[body] = bq_per_voxel(0.02, voxel_mm)
    for _ in range(10):
        r = int(rng.integers(5, 21))
        cc = [rng.uniform(0.3, 0.7) * s for s in shape]
        _fill_sphere(act0, cc, r, bq_per_voxel(float(rng.uniform(1, 5)), voxel_mm))
    lungs = [((sx / 2, 0.32 * sy, 0.7 * sz), (0.2 * sx, 0.12 * sy, 0.2 * sz)),
             ((sx / 2, 0.68 * sy, 0.7 * sz), (0.2 * sx, 0.12 * sy, 0.2 * sz))]
    rho = _body_density(shape, lungs, ((0.75 * sx, sy / 2), 0.04 * sx), soft=1.04, lung_rho=0.26)
    teff = float(rng.uniform(40, 120))
    lam = math.log(2.0) / teff
    acts = [(act0 * np.float32(math.exp(-lam * t))).astype(np.float32) for t in TORSO_TIMEPOINTS_H]
    return Phantom("torso", voxel_mm, acts, rho, list(TORSO_TIMEPOINTS_H), "Lu177", (61, 61, 61),
                   {"t_eff_h": teff})


def synthetic_kernel(voxel_mm: float, grid: Sequence[int], length_mm: float = 2.5, core_mm2: float = 0.25):
    """Isotropic k(r) ~ exp(-r/2.5 mm)/(r^2 + 0.25) normalised to sum 1, centre g//2."""
    c = [g // 2 for g in grid]
    x, y, z = np.ogrid[:grid[0], :grid[1], :grid[2]]
    r = np.sqrt(((x - c[0]) * voxel_mm) ** 2 + ((y - c[1]) * voxel_mm) ** 2 + ((z - c[2]) * voxel_mm) ** 2)
    k = np.exp(-r / length_mm) / (r ** 2 + core_mm2)
    return k / k.sum()


def ct_hu_phantom(seed: int = 6, shape=(512, 512, 400), metal: bool = True) -> np.ndarray:
    """Synthetic torso CT in HU (float32): air -1000 outside the body, soft
    tissue 40 +- 15, fat -100, lungs -850 +- 40, spine 700 +- 150 and, with
    `metal`, two dense implants of 3000 HU (a hip-like sphere and a thin rod)."""
    rng = np.random.default_rng(seed)
    sx, sy, sz = shape
    ct = np.full(shape, -1000.0, dtype=np.float32)
    body = _ellipsoid(shape, (sx / 2, sy / 2, sz / 2), (0.45 * sx, 0.35 * sy, 0.6 * sz))
    ct[body] = 40.0
    fat = body & ~_ellipsoid(shape, (sx / 2, sy / 2, sz / 2), (0.41 * sx, 0.31 * sy, 0.58 * sz))
    ct[fat] = -100.0
    for cy in (0.32, 0.68):
        ct[_ellipsoid(shape, (sx / 2, cy * sy, 0.7 * sz), (0.2 * sx, 0.12 * sy, 0.2 * sz))] = -850.0
    x, y = np.ogrid[:sx, :sy]
    spine = ((x - 0.75 * sx) ** 2 + (y - sy / 2) ** 2) <= (0.04 * sx) ** 2
    ct[spine[:, :, None] & body] = 700.0
    noise = rng.normal(0.0, 1.0, size=shape).astype(np.float32)
    scale = np.where(ct == -850.0, 40.0, np.where(ct == 700.0, 150.0, np.where(ct == -1000.0, 5.0, 15.0)))
    ct += (noise * scale).astype(np.float32)
    if metal:
        _fill_sphere(ct, (0.5 * sx, 0.3 * sy, 0.2 * sz), max(2, int(0.03 * sx)), 3000.0)
        ct[int(0.6 * sx), int(0.6 * sy), int(0.1 * sz):int(0.5 * sz)] = 3000.0
    return ct
