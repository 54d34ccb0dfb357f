"""GPU tests of the HU -> tissue kernel (include/pvd_tissue.h) against oracle/tissue.py.

Tolerances: the metal fill is an FP64 sum rounded to FP32 (<= 0.5 ulp of the
FP32 result, 2.4e-4 HU at 3000 HU); masks and densities are exact except for
voxels whose filled value lies within 1e-3 HU of a class or material boundary
(a few hundredths of a percent of these noisy inputs; the test counts them)."""
import sys

import numpy as np
import pytest

from oracle import tissue as ot

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[gpu-case]", *a, file=sys.stderr, flush=True)


def _near_boundary(v, tol=1e-3):
    from pyvoxeldosimetry_amd.tissue import HU_CLASSES, HU_MATERIALS
    edges = [e for lo_hi in HU_CLASSES.values() for e in lo_hi] + [m[0] for m in HU_MATERIALS]
    return np.any(np.abs(v[..., None] - np.array(edges)) < tol, axis=-1)


def _check(ct, handle=True):
    import torch
    from pyvoxeldosimetry_amd.tissue import HU_CLASSES, HU_MATERIALS, TissueComposition
    tc = TissueComposition()
    corr, bits, dens = tc.map(torch.from_numpy(ct).cuda(), handle_artifacts=handle)
    corr, bits, dens = corr.cpu().numpy(), bits.cpu().numpy(), dens.cpu().numpy()
    exp = ot.handle_artifacts(ct) if handle else ct.astype(np.float64)
    fin = np.isfinite(exp)
    assert np.array_equal(np.isnan(corr), np.isnan(exp))
    np.testing.assert_allclose(corr[fin], exp[fin], rtol=1e-6, atol=5e-4)
    skip = _near_boundary(exp)
    comp = ot.calculate_composition(ct, handle=handle)
    for c, name in enumerate(HU_CLASSES):
        got = ((bits >> c) & 1).astype(np.float64)
        assert np.array_equal(got[~skip], comp[name][~skip]), name
    de = ot.density_from_hu(exp, HU_MATERIALS)
    np.testing.assert_array_equal(np.isnan(dens), np.isnan(de))
    ok = ~skip & ~np.isnan(de)
    np.testing.assert_array_equal(dens[ok], de[ok].astype(np.float32))
    return int(skip.sum()), int((ct > 2000).sum())


@pytest.mark.parametrize("shape", [(64, 64, 48), (37, 29, 23)])
def test_tissue_map_matches_oracle(shape):
    from pyvoxeldosimetry_amd.phantoms import ct_hu_phantom
    _say("tissue map", shape)
    ct = ct_hu_phantom(seed=7, shape=shape)
    ct[0, 0, 0] = ct[-1, -1, -1] = ct[0, -1, 3] = 2999.0   # metal on corners and edges (reflected windows)
    ct[5, 5, 5] = np.nan
    skipped, n_metal = _check(ct)
    _say(f"metal voxels {n_metal}, near-boundary voxels skipped {skipped}")
    assert n_metal > 10 and skipped < 1e-3 * ct.size


def test_tissue_map_grid_smaller_than_filter_and_no_fill():
    _say("tissue map tiny grid")
    rng = np.random.default_rng(11)
    ct = rng.uniform(-1000, 1500, size=(3, 2, 5)).astype(np.float32)
    ct[1, 0, 2] = ct[2, 1, 4] = 2600.0
    _check(ct)
    _check(ct, handle=False)


def test_reference_api_and_density():
    from pyvoxeldosimetry_amd.phantoms import ct_hu_phantom
    from pyvoxeldosimetry_amd.tissue import HU_MATERIALS, TissueComposition, density_from_hu
    _say("tissue drop-in API")
    ct = ct_hu_phantom(seed=8, shape=(40, 36, 32)).astype(np.float64)
    comp = TissueComposition().calculate_composition(ct)
    exp = ot.calculate_composition(ct)
    assert list(comp) == list(exp)
    for k in exp:
        assert comp[k].dtype == np.float64 and np.array_equal(comp[k], exp[k]), k
    np.testing.assert_allclose(TissueComposition()._handle_artifacts(ct), ot.handle_artifacts(ct), rtol=1e-6,
                               atol=5e-4)
    d = density_from_hu(ct)
    np.testing.assert_array_equal(d, ot.density_from_hu(ot.handle_artifacts(ct), HU_MATERIALS).astype(np.float32))


def test_density_feeds_dose_calculator():
    """The producer's density map drives the dose epilogue: dose = conv * rho_w / rho."""
    import torch
    from oracle import conv, dpk
    from pyvoxeldosimetry_amd import KernelConvolutionCalculator
    from pyvoxeldosimetry_amd.phantoms import ct_hu_phantom
    from pyvoxeldosimetry_amd.tissue import density_from_hu
    _say("density -> dose")
    ct = ct_hu_phantom(seed=9, shape=(48, 40, 36))
    rho = density_from_hu(ct)
    act = np.zeros(ct.shape, np.float32)
    act[20:28, 15:25, 10:20] = 1e5
    calc = KernelConvolutionCalculator("Y90", "water", 2.0, {"kernel_grid": (11, 11, 11), "fix_kernel_centre": True})
    got = calc.calculate_dose_rate(act, (2.0, 2.0, 2.0), density_map=rho)
    k = dpk.y90_kernel(2.0, (11, 11, 11), fix_centre=True)
    exp = conv.epilogue(conv.ref_conv(act.astype(np.float64), k), rho)
    assert conv.error_metrics(got, exp)["max_rel"] < 1e-5
    torch.cuda.synchronize()


@pytest.mark.parametrize("shape", [(20, 18, 24), (9, 11, 13)])
def test_metal_fill_other_sigma_and_a_long_rod(shape):
    """sigma 0.6 (radius 2, the run-time-radius kernel) and a metal rod along z."""
    import torch
    from pyvoxeldosimetry_amd.tissue import TissueComposition
    _say("tissue sigma 0.6", shape)
    rng = np.random.default_rng(21)
    ct = rng.uniform(-900, 1500, size=shape).astype(np.float32)
    ct[shape[0] // 2, shape[1] // 3, :] = 3100.0            # rod through the whole z extent
    ct[0, :, 0] = 2500.0
    corr, _, _ = TissueComposition({"sigma": 0.6}).map(torch.from_numpy(ct).cuda())
    exp = ot.handle_artifacts(ct, sigma=0.6)
    np.testing.assert_allclose(corr.cpu().numpy(), exp, rtol=1e-6, atol=5e-4)


def test_ct_phantom_density_is_its_hu_map_through_the_device_producer():
    """The CT workload's density map: its synthetic CT through pvd_tissue_map on the GPU
    equals the phantom's table lookup bit for bit."""
    import torch
    from pyvoxeldosimetry_amd.phantoms import ct_phantom
    from pyvoxeldosimetry_amd.tissue import density_from_hu
    p = ct_phantom()
    rho = density_from_hu(torch.from_numpy(p.meta["hu"]).cuda())
    assert torch.equal(rho.cpu(), torch.from_numpy(p.density))
