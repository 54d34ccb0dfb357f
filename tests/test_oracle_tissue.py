"""Pins of the HU -> tissue oracle (oracle/tissue.py) on the CPU."""
import json
import os

import numpy as np
import pytest

from oracle import tissue as ot
from tests.conftest import ROOT


def test_no_metal_is_identity():
    rng = np.random.default_rng(0)
    ct = rng.uniform(-1000, 1999, size=(9, 7, 11))
    np.testing.assert_array_equal(ot.handle_artifacts(ct), ct)


def test_single_metal_voxel_in_constant_field_closed_form():
    c = 40.0
    ct = np.full((15, 15, 15), c)
    ct[7, 7, 7] = 3000.0
    w = ot.gaussian_weights(1.0)
    got = ot.handle_artifacts(ct)[7, 7, 7]
    assert got == pytest.approx(c * (1.0 - w[len(w) // 2] ** 3), rel=1e-13)
    assert len(w) == 9                                     # radius int(4*1 + 0.5) = 4


@pytest.mark.parametrize("shape", [(6, 7, 8), (3, 2, 5), (1, 4, 9)])
def test_fill_matches_brute_force_with_reflection(shape):
    rng = np.random.default_rng(sum(shape))
    ct = rng.uniform(-800, 1500, size=shape)
    metal = [tuple(int(rng.integers(0, s)) for s in shape) for _ in range(3)] + [(0, 0, 0),
                                                                               tuple(s - 1 for s in shape)]
    centre = tuple(s // 2 for s in shape)
    metal = [m for m in metal if m != centre]
    for m in metal:
        ct[m] = 2500.0 + rng.uniform(0, 500)
    ct[centre] = np.nan                                     # non-metal NaN: stays NaN, smooths as 0
    out = ot.handle_artifacts(ct)
    for m in set(metal):
        assert out[m] == pytest.approx(ot.brute_fill_at(ct, m), rel=1e-12, abs=1e-9)
    assert np.isnan(out[centre])


def test_composition_ranges_are_inclusive():
    ct = np.array([-1000, -900, -899.5, -500, -100, -10, 10, 100, 300, 3000, 3000.5, 250], float).reshape(2, 2, 3)
    comp = ot.calculate_composition(ct, handle=False)
    flat = {k: v.reshape(-1) for k, v in comp.items()}
    assert flat["air"].tolist()[:2] == [1.0, 1.0] and flat["lung"][1] == 1.0
    assert flat["soft_tissue"][4] == 1.0 and flat["water"][5] == 1.0 and flat["water"][6] == 1.0
    assert flat["bone"][8] == 1.0 and flat["bone"][9] == 1.0 and flat["bone"][10] == 0.0
    assert sum(v[11] for v in flat.values()) == 0.0        # 250 HU is in no class (gap 100..300)
    assert all(v.dtype == np.float64 for v in comp.values())


def test_material_table_matches_reference_data_files():
    """HU_MATERIALS against HU_to_material.txt and GateMaterials.db (fixture)."""
    from pyvoxeldosimetry_amd.tissue import HU_MATERIALS
    ref = json.load(open(os.path.join(ROOT, "tests", "golden", "hu_materials.json")))
    assert [m[0] for m in HU_MATERIALS] == [r[0] for r in ref["ranges"]]
    assert [m[1] for m in HU_MATERIALS] == [r[2] for r in ref["ranges"]]
    # the file's integer ranges tile [-1050, 3000] without gaps
    for a, b in zip(ref["ranges"], ref["ranges"][1:]):
        assert b[0] == a[1] + 1
    declared = {"Tissue": 1.04, "Bone": 1.85, "Tooth": 2.75}
    for lo, name, rho in HU_MATERIALS:
        exp = ref["densities"][name]
        assert rho == pytest.approx(exp if exp is not None else declared[name], rel=1e-12)


def test_density_lookup_boundaries():
    from pyvoxeldosimetry_amd.tissue import HU_MATERIALS
    lo = [m[0] for m in HU_MATERIALS]
    rho = [m[2] for m in HU_MATERIALS]
    hu = np.array([lo[0] - 100] + lo + [l - 0.5 for l in lo[1:]] + [5000.0, np.nan])
    d = ot.density_from_hu(hu, HU_MATERIALS)
    n = len(lo)
    assert d[0] == rho[0]
    assert d[1:1 + n].tolist() == rho
    assert d[1 + n:2 * n].tolist() == rho[:-1]
    assert d[-2] == rho[-1] and np.isnan(d[-1])
