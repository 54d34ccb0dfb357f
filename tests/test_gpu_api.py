"""GPU tests of the drop-in API (calculator / facade / sampler) against the oracle."""
import sys

import numpy as np
import pytest

from oracle import conv, dpk

pytestmark = pytest.mark.gpu

MAX_REL = 1e-5


def _say(*a):
    print("[gpu-case]", *a, file=sys.stderr, flush=True)


def rel(got, ref):
    return conv.error_metrics(got, ref)["max_rel"]


def test_reference_example_48cubed_default_64_kernel_is_nan():
    """single_timepoint_y90_physical_decay.py: 48^3, r=8, 2e6 Bq, Y-90 at 1 mm with the
    reference's 64^3 kernel (cropped to 48^3) -> the NaN centre voxel makes all dose NaN."""
    from pyvoxeldosimetry_amd import DoseCalculator
    from pyvoxeldosimetry_amd.phantoms import create_spherical_activity
    _say("reference example Y90 48^3")
    act = create_spherical_activity((48, 48, 48), (24, 24, 24), 8, 2e6)
    calc = DoseCalculator("Y90", "kernel", {"kernel_resolution": 1.0, "tissue_composition": "water",
                                            "half_life": 64.1, "time_units": "hours"})
    res = calc.calculate_dose(activity_maps=[act], time_points=[2.0], voxel_size=(1.0, 1.0, 1.0))
    assert res.absorbed_dose is None and res.metadata == {"mode": "single_timepoint"}
    assert np.isnan(res.dose_rate_maps[0]).all()
    assert calc.calculator.kernel.shape == (64, 64, 64)
    assert np.isnan(calc.calculator.kernel).sum() == 1


def test_calculator_cropped_default_kernel_fixed_centre():
    from pyvoxeldosimetry_amd import KernelConvolutionCalculator
    from pyvoxeldosimetry_amd.phantoms import create_spherical_activity
    _say("calculator 48^3 with cropped 64^3 kernel")
    act = create_spherical_activity((48, 48, 48), (24, 24, 24), 8, 2e6)
    calc = KernelConvolutionCalculator("Y90", "water", 1.0, {"fix_kernel_centre": True, "out_dtype": "float64"})
    got = calc.calculate_dose_rate(act, (1.0, 1.0, 1.0))
    assert got.dtype == np.float64
    exp = conv.ref_conv(act, dpk.y90_kernel(1.0, (64, 64, 64), fix_centre=True))
    assert rel(got, exp) <= MAX_REL
    strict = KernelConvolutionCalculator("Y90", "water", 1.0, {"fix_kernel_centre": True, "resample": False})
    with pytest.raises(ValueError, match="kernel_resolution"):
        strict.calculate_dose_rate(act, (2.0, 2.0, 2.0))


def test_calculator_absorbed_dose_time_series_lu177():
    """time_integrated_dose.py workload: 48^3 sphere, Lu-177, 6 timepoints."""
    from pyvoxeldosimetry_amd import KernelConvolutionCalculator
    from pyvoxeldosimetry_amd.phantoms import create_spherical_activity, create_time_series
    _say("absorbed dose Lu177 48^3 T=6")
    init = create_spherical_activity((48, 48, 48), (24, 24, 24), 10, 1000.0)
    t = [0, 24, 48, 72, 96, 120]
    maps = create_time_series(init, 161.52, t)
    calc = KernelConvolutionCalculator("Lu177", "water", 1.0, {"kernel_grid": (31, 31, 31)})
    got = calc.calculate_absorbed_dose(maps, t, (1.0, 1.0, 1.0))
    exp = conv.ref_absorbed_dose(maps, t, dpk.lu177_kernel(1.0, (31, 31, 31)))
    assert rel(got, exp) <= MAX_REL


def test_facade_modes_agree_with_oracle():
    import torch
    from pyvoxeldosimetry_amd import DoseCalculator
    from pyvoxeldosimetry_amd.phantoms import create_spherical_activity, create_time_series
    _say("facade modes")
    init = create_spherical_activity((40, 36, 32), (20, 18, 16), 7, 5000.0)
    t = [1.0, 4.0, 24.0, 72.0]
    maps = create_time_series(init, 64.1, t)
    rho = np.random.default_rng(0).uniform(0.3, 1.9, size=init.shape).astype(np.float32)
    cfg = {"kernel_resolution": 2.0, "half_life": 64.1, "kernel_grid": (11, 11, 11), "fix_kernel_centre": True}
    calc = DoseCalculator("Y90", "kernel", cfg)
    k = dpk.y90_kernel(2.0, (11, 11, 11), fix_centre=True)
    exp = conv.epilogue(conv.ref_absorbed_dose(maps, t, k), rho)
    r1 = calc.calculate_dose(maps, t, (2.0, 2.0, 2.0), tissue_densities=rho, integration_mode="activity")
    assert r1.metadata["mode"] == "multi_timepoint_activity"
    assert rel(r1.absorbed_dose, exp) <= MAX_REL
    r2 = calc.calculate_dose(maps, t, (2.0, 2.0, 2.0), tissue_densities=rho, integration_mode="dose_rate")
    assert len(r2.dose_rate_maps) == 4
    assert rel(r2.absorbed_dose, exp) <= MAX_REL
    for i in range(4):
        assert rel(r2.dose_rate_maps[i], conv.epilogue(conv.ref_conv(maps[i], k), rho)) <= MAX_REL
    # dose_rate mode with a decay tail to 10 half-lives
    lam = np.log(2) / 64.1
    r3 = calc.calculate_dose(maps, t, (2.0, 2.0, 2.0), integration_mode="dose_rate", integration_limit=641.0)
    tail = conv.ref_conv(maps[-1], k) * (1 - np.exp(-lam * (641.0 - 72.0))) / lam * 3600.0
    assert rel(r3.absorbed_dose, conv.ref_absorbed_dose(maps, t, k) + tail) <= MAX_REL
    # accumulated activity input (Bq h)
    acc = np.trapezoid(np.stack(maps), x=t, axis=0)
    r4 = calc.calculate_dose(accumulated_activity=acc, voxel_size=(2.0, 2.0, 2.0))
    assert rel(r4.absorbed_dose, conv.ref_absorbed_dose(maps, t, k)) <= MAX_REL
    # single timepoint with physical decay
    cfg2 = dict(cfg, physical_decay=True)
    r5 = DoseCalculator("Y90", "kernel", cfg2).calculate_dose([maps[0]], [2.0], (2.0, 2.0, 2.0))
    assert rel(r5.dose_rate_maps[0], conv.ref_conv(maps[0], k) * 2 ** (-2.0 / 64.1)) <= MAX_REL
    # torch tensors in -> torch tensors out, on the device
    tt = torch.from_numpy(maps[0].astype(np.float32)).cuda()
    out = calc.calculator.calculate_dose_rate(tt, (2.0, 2.0, 2.0))
    assert isinstance(out, torch.Tensor) and out.is_cuda
    with pytest.raises(ValueError, match="Unknown integration_mode"):
        calc.calculate_dose(maps, t, (2.0, 2.0, 2.0), integration_mode="bogus")
    with pytest.raises(ValueError, match="must match"):
        calc.calculate_dose(maps, t[:3], (2.0, 2.0, 2.0))


def test_activity_sampler_integrate_activity():
    from pyvoxeldosimetry_amd import ActivitySampler
    _say("activity sampler")
    rng = np.random.default_rng(4)
    maps = [rng.random((10, 11, 12)).astype(np.float32) for _ in range(4)]
    t = [0.0, 2.0, 7.0, 30.0]
    got = ActivitySampler(6.0).integrate_activity(maps, t)
    np.testing.assert_allclose(got, conv.activity_trapezoid(maps, t), rtol=1e-6)


def test_autotuned_plan_matches(tmp_path, monkeypatch):
    _say("autotune")
    monkeypatch.setenv("PVD_AUTOTUNE_CACHE", str(tmp_path / "tune.json"))
    import importlib
    from pyvoxeldosimetry_amd import autotune
    importlib.reload(autotune)
    from pyvoxeldosimetry_amd import KernelConvolutionCalculator
    import pyvoxeldosimetry_amd.core.kernel_convolution as kc
    rng = np.random.default_rng(8)
    a = rng.random((50, 47, 41))
    calc = KernelConvolutionCalculator("Lu177", "water", 2.0, {"kernel_grid": (9, 9, 9), "autotune": True})
    got = calc.calculate_dose_rate(a, (2.0, 2.0, 2.0))
    exp = conv.ref_conv(a, dpk.lu177_kernel(2.0, (9, 9, 9)))
    assert rel(got, exp) <= MAX_REL
    assert (tmp_path / "tune.json").exists()
    import json
    entry = next(iter(json.load(open(tmp_path / "tune.json")).values()))
    if entry["algo"] == "fft":
        assert len(entry["tile"]) == 3 and entry["max_radix"] in (8, 16) and entry["stockham"] in (True, False)
    # a second calculator takes the cached choice and gives the same bits
    calc2 = KernelConvolutionCalculator("Lu177", "water", 2.0, {"kernel_grid": (9, 9, 9), "autotune": True})
    np.testing.assert_array_equal(calc2.calculate_dose_rate(a, (2.0, 2.0, 2.0)), got)
    del kc
