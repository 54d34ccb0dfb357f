"""Write the golden fixtures of tests/golden/ from the FP64 oracle.

    python tests/golden/make_golden.py

Each case stores its inputs and the oracle's FP64 output.  The oracle itself
is pinned by tests/test_oracle.py (brute force in Python and in C, closed
forms, library routines); these fixtures freeze its outputs so the GPU tests
compare against data, and the CPU suite re-derives them to catch drift.
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import conv, dpk  # noqa: E402

# (name, shape, kernel source, kernel grid, mode, extra)
CONV_CASES = [
    ("ref_small_aniso", (8, 9, 10), "rand", (3, 5, 4), "reference", {}),
    ("ref_crop", (6, 7, 5), "rand", (9, 7, 5), "reference", {}),
    ("ref_nonsmooth", (11, 13, 10), "rand", (3, 3, 3), "reference", {}),
    ("ref_prime", (17, 19, 23), "rand", (5, 5, 5), "reference", {}),
    ("ref_y90_24", (24, 24, 24), "y90fix", (7, 7, 7), "reference", {"voxel": 4.42}),
    ("ref_lu177_even_kernel", (24, 24, 20), "lu177", (8, 8, 8), "reference", {"voxel": 2.0}),
    ("ref_density_scale", (16, 12, 20), "rand", (5, 5, 5), "reference", {"density": True, "scale": 0.75}),
    ("lin_small", (12, 10, 14), "rand", (5, 5, 5), "linear", {}),
    ("lin_big_kernel", (6, 6, 6), "rand", (15, 4, 3), "linear", {}),
    ("lin_even_kernel", (10, 11, 9), "rand", (4, 6, 2), "linear", {}),
    ("lin_density", (20, 18, 16), "lu177", (9, 9, 9), "linear", {"voxel": 2.0, "density": True}),
    ("ts_trapezoid", (16, 16, 16), "rand", (5, 5, 5), "reference", {"times": [0.0, 24.0, 48.0, 72.0]}),
    ("ts_uneven", (14, 12, 10), "y90fix", (5, 5, 5), "linear", {"times": [1.0, 4.0, 24.0], "voxel": 4.42}),
    ("ref_y90_nan", (12, 12, 12), "y90raw", (5, 5, 5), "reference", {"voxel": 1.0}),
    ("ref_single_voxel", (1, 1, 1), "rand", (3, 3, 3), "reference", {}),
    ("lin_thin", (1, 9, 33), "rand", (3, 3, 7), "linear", {}),
]

DPK_CASES = [
    ("dpk_y90_water_fix", "y90", "water", 4.42, (11, 11, 11), True),
    ("dpk_y90_bone_raw", "y90", "bone", 1.0, (9, 9, 9), False),
    ("dpk_y90_lung_even", "y90", "lung", 2.0, (8, 10, 6), True),
    ("dpk_lu177_water", "lu177", "water", 2.0, (21, 21, 21), None),
    ("dpk_lu177_lung", "lu177", "lung", 0.98, (15, 15, 15), None),
    ("dpk_lu177_unknown_tissue", "lu177", "liver", 1.0, (7, 7, 7), None),
    ("dpk_ga68_water_fix", "ga68", "water", 2.0, (25, 25, 25), True),
    ("dpk_ga68_lung_raw", "ga68", "lung", 1.0, (15, 15, 15), False),
    ("dpk_ga68_bone_even", "ga68", "bone", 0.98, (10, 12, 8), True),
]


def make_kernel(src, grid, voxel, rng):
    if src == "rand":
        return rng.random(grid)
    if src == "y90fix":
        return dpk.y90_kernel(voxel, grid, "water", fix_centre=True)
    if src == "y90raw":
        return dpk.y90_kernel(voxel, grid, "water", fix_centre=False)
    if src == "lu177":
        return dpk.lu177_kernel(voxel, grid, "water")
    raise ValueError(src)


def build_cases():
    rng = np.random.default_rng(20250202)
    out = {}
    for name, shape, src, kgrid, mode, ex in CONV_CASES:
        k = make_kernel(src, kgrid, ex.get("voxel", 1.0), rng)
        times = ex.get("times")
        nmaps = len(times) if times else 1
        acts = [rng.random(shape).astype(np.float32) * 100.0 for _ in range(nmaps)]
        dens = (rng.uniform(0.2, 2.0, size=shape).astype(np.float32) if ex.get("density") else None)
        if dens is not None:
            dens.flat[0] = 0.0          # rho <= 0 voxel -> dose 0
        scale = ex.get("scale", 1.0)
        expect = conv.ref_dose([a.astype(np.float64) for a in acts], k, time_points=times, density=dens,
                               scale=scale, mode=mode)
        case = {"kernel": k, "expect": expect, "mode": np.array(mode), "scale": np.array(scale),
                "acts": np.stack(acts)}
        if times:
            case["times"] = np.array(times, dtype=np.float64)
        if dens is not None:
            case["density"] = dens
        out[name] = case
    for name, nuc, tissue, voxel, grid, fix in DPK_CASES:
        if nuc == "y90":
            k = dpk.y90_kernel(voxel, grid, tissue, fix_centre=fix)
        elif nuc == "ga68":
            k = dpk.ga68_kernel(voxel, grid, tissue, fix_centre=fix)
        else:
            k = dpk.lu177_kernel(voxel, grid, tissue)
        out[name] = {"expect": k, "nuclide": np.array({"y90": "Y90", "ga68": "Ga68", "lu177": "Lu177"}[nuc]),
                     "tissue": np.array(tissue), "voxel": np.array(voxel), "grid": np.array(grid),
                     "fix_centre": np.array(bool(fix))}
    return out


def main():
    cases = build_cases()
    path = os.path.join(HERE, "golden_v1.npz")
    flat = {f"{c}/{k}": v for c, d in cases.items() for k, v in d.items()}
    np.savez_compressed(path, **flat)
    print(f"wrote {path}: {len(cases)} cases, {os.path.getsize(path)} bytes")


if __name__ == "__main__":
    main()
