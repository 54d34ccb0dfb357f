"""CPU tests of the host logic and the C ABI surface (no GPU calls)."""
import ctypes
import os
import re

import numpy as np
import pytest

from tests.conftest import ROOT

HEADERS = [os.path.join(ROOT, "include", h) for h in sorted(os.listdir(os.path.join(ROOT, "include")))
           if h.endswith(".h")]


def header_functions():
    names = set()
    for h in HEADERS:
        text = re.sub(r"/\*.*?\*/", "", open(h).read(), flags=re.S)
        names |= set(re.findall(r"\b(pvd_[a-z0-9_]+)\s*\(", text))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    from pyvoxeldosimetry_amd import _native
    lib = ctypes.CDLL(_native.LIB_PATH)
    missing = [f for f in header_functions() if not hasattr(lib, f)]
    assert not missing, missing


def test_binding_types_every_declared_symbol():
    from pyvoxeldosimetry_amd import _native
    assert sorted(_native.SIGNATURES) == header_functions()


def test_binding_loads_and_reports_version_without_gpu():
    from pyvoxeldosimetry_amd import _native
    lib = _native.load()
    assert lib.pvd_version() >= 100


def test_struct_layouts_match_header():
    from pyvoxeldosimetry_amd import _native as nat
    assert ctypes.sizeof(nat.PlanParams) == 4 * 16
    assert ctypes.sizeof(nat.DpkTerm) == 24
    # 9 ints, padding, size_t, 17 ints (tile, max_radix, fft_path, block, grid), tail padding
    assert ctypes.sizeof(nat.PlanInfo) == 9 * 4 + 4 + 8 + 17 * 4 + 4
    # pvd_tissue_params: 2 ints + 2x8 + 2x32 floats + int + float, then two doubles
    assert ctypes.sizeof(nat.TissueParams) == 4 * (2 + 16 + 64 + 2) + 16
    assert nat.TissueParams.sigma.offset == 4 * (2 + 16 + 64 + 2)


def test_product_path_fails_loudly_without_gpu():
    import torch
    if torch.cuda.is_available():
        pytest.skip("GPU present")
    from pyvoxeldosimetry_amd import ConvPlan
    with pytest.raises(RuntimeError, match="GPU"):
        ConvPlan((8, 8, 8), (3, 3, 3))


def test_missing_library_raises(tmp_path):
    from pyvoxeldosimetry_amd import _native
    with pytest.raises(_native.NativeLibraryMissing):
        _native.load(str(tmp_path / "nope.so"))


# --- reference error behaviour (raised before any device work) -------------------------
def test_non_string_radionuclide_is_type_error():
    from pyvoxeldosimetry_amd import KernelConvolutionCalculator
    with pytest.raises(TypeError, match="Radionuclide must be a string"):
        KernelConvolutionCalculator(90, "water")


def test_unknown_nuclide_is_value_error():
    from pyvoxeldosimetry_amd import KernelFactory
    with pytest.raises(ValueError, match="Unsupported nuclide: F18"):
        KernelFactory().get_kernel("F18", "water")


def test_dose_calculator_rejects_unknown_method():
    from pyvoxeldosimetry_amd import DoseCalculator
    with pytest.raises(ValueError, match="Unsupported calculation method"):
        DoseCalculator("Y90", method="magic")


def test_activity_sampler_checks():
    from pyvoxeldosimetry_amd import ActivitySampler
    with pytest.raises(ValueError, match="Half-life must be positive"):
        ActivitySampler(0.0)
    with pytest.raises(ValueError, match="Invalid time units"):
        ActivitySampler(1.0, units="days")
    s = ActivitySampler(64.1)
    np.testing.assert_allclose(s.sample_timepoints(0, 10, 5), np.linspace(0, 10, 5))


def test_trapezoid_weights_match_numpy():
    from pyvoxeldosimetry_amd.core.activity_sampler import trapezoid_weights
    t = [1.0, 4.0, 24.0, 48.0, 96.0, 168.0]
    x = np.random.default_rng(0).random(len(t))
    assert np.dot(trapezoid_weights(t), x) == pytest.approx(np.trapezoid(x, t), rel=1e-14)
    assert trapezoid_weights([5.0]) == [0.0]


# --- workloads ---------------------------------------------------------------------------
def test_spherical_activity_matches_loop_definition():
    from pyvoxeldosimetry_amd.phantoms import create_spherical_activity
    a = create_spherical_activity((12, 11, 10), (5, 6, 4), 3, 7.0)
    exp = np.zeros((12, 11, 10))
    for x in range(12):
        for y in range(11):
            for z in range(10):
                if (x - 5) ** 2 + (y - 6) ** 2 + (z - 4) ** 2 <= 9:
                    exp[x, y, z] = 7.0
    np.testing.assert_array_equal(a, exp)


def test_small_phantom_is_seeded_and_sized():
    from pyvoxeldosimetry_amd.phantoms import bq_per_voxel, small_phantom
    p1, p2 = small_phantom(1), small_phantom(1)
    np.testing.assert_array_equal(p1.activity[0], p2.activity[0])
    assert p1.shape == (64, 64, 64)
    assert bq_per_voxel(1.0, 4.42) == pytest.approx(86350.888, rel=1e-6)
    inside = p1.activity[0][p1.activity[0] > 0]
    assert abs(inside.mean() / bq_per_voxel(1.0, 4.42) - 1) < 0.01
    assert 0.03 < inside.std() / inside.mean() < 0.07      # 5 % Poisson noise


def test_batch_specs_ranges():
    from pyvoxeldosimetry_amd.phantoms import batch_specs
    specs = batch_specs()
    assert len(specs) == 48
    for s in specs:
        assert 96 <= s.shape[0] <= 384 and 96 <= s.shape[1] <= 384 and 96 <= s.shape[2] <= 320
        assert 20 <= s.t_eff_h <= 150
        assert (s.nuclide, s.kernel_grid) in (("Y90", (11, 11, 11)), ("Lu177", (31, 31, 31)))


def test_ct_phantom_density_comes_from_its_hu_map():
    """The CT phantom's density is its synthetic HU map through the HU -> material
    table (checked against the oracle's restatement of the reference's table): soft
    tissue 1.04, lung 0.26, bone 1.85 g/cc."""
    from oracle import tissue as ot
    from pyvoxeldosimetry_amd.phantoms import ct_phantom
    from pyvoxeldosimetry_amd.tissue.composition import HU_MATERIALS
    p = ct_phantom()
    hu = p.meta["hu"]
    assert hu.shape == p.shape and hu.dtype == np.float32
    np.testing.assert_array_equal(p.density, ot.density_from_hu(hu, HU_MATERIALS).astype(np.float32))
    vals = set(np.unique(p.density).astype(np.float64).round(3).tolist())
    assert vals == {0.26, 1.04, 1.85}
    assert p.activity[0].max() > 10 * np.median(p.activity[0][p.activity[0] > 0])


def test_autotune_candidate_lengths_do_not_alias():
    from pyvoxeldosimetry_amd.autotune import candidate_lengths, _smooth
    for n, k, mode in [(97, 11, "reference"), (256, 31, "reference"), (200, 31, "linear"), (5, 64, "reference")]:
        for p in candidate_lengths(n, k, mode):
            assert _smooth(p)
            if mode == "reference":
                assert p == n or p >= n + min(k, n) - 1
            else:
                assert p >= n + k // 2


def test_shipped_cnn_weights_are_well_formed():
    """The pretrained weights are data only (safetensors), of the network's size,
    finite, and zero where the architecture fixes them (layer 0 reads 3 input
    channels, layer 5 writes one)."""
    from safetensors import safe_open

    from pyvoxeldosimetry_amd.learned.cnn import CIN
    from pyvoxeldosimetry_amd.learned.pipeline import PRETRAINED, WEIGHTS_DIR
    for name in PRETRAINED.values():
        with safe_open(os.path.join(WEIGHTS_DIR, name), framework="np") as f:
            meta = f.metadata()
            w = f.get_tensor("master")
        assert w.dtype == np.float32
        assert w.size == sum(27 * 32 * c + 32 for c in CIN)
        assert np.isfinite(w).all()
        w0 = w[:27 * 32 * 16].reshape(27, 32, 16)
        assert not w0[:, :, 3:].any() and w0[:, :, :3].any()
        off5 = sum(27 * 32 * c + 32 for c in CIN[:5])
        w5 = w[off5:off5 + 27 * 32 * 32].reshape(27, 32, 32)
        assert not w5[:, 1:, :].any() and w5[:, 0, :].any()
        assert {"nuclide", "voxel_mm", "train_seeds", "command"} <= set(meta)


def test_struct_layouts_match_the_c_compiler(tmp_path):
    """sizeof/offsetof of every ABI struct as gcc lays it out, and the header's constants,
    == the ctypes mirror."""
    import shutil
    import subprocess
    from pyvoxeldosimetry_amd import _native as nat
    if shutil.which("gcc") is None:
        pytest.skip("gcc not available")
    fields = {"pvd_plan_params": (nat.PlanParams, ["flags", "tile", "max_radix"]),
              "pvd_plan_info": (nat.PlanInfo, ["workspace_bytes", "tile", "max_radix", "fft_path", "block", "grid"]),
              "pvd_dpk_term": (nat.DpkTerm, ["param"]), "pvd_mc_params": (nat.McParams, ["active", "n_active", "line_energy", "brems_z", "mu_en", "mu_log_e_hi"]),
              "pvd_tissue_params": (nat.TissueParams, ["sigma", "truncate"]),
              "pvd_tac_params": (nat.TacParams, ["tol", "fail_params", "t_int"])}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "pvd.h"', '#include "pvd_learned.h"',
             '#include "pvd_tissue.h"', '#include "pvd_tac.h"', "int main(void) {"]
    for name, (_, fl) in fields.items():
        lines.append(f'printf("{name} %zu\\n", sizeof({name}));')
        for f in fl:
            lines.append(f'printf("{name}.{f} %zu\\n", offsetof({name}, {f}));')
    consts = ["PVD_CNN_EPI_RELU", "PVD_CNN_EPI_RESID_RELU", "PVD_CNN_EPI_LINEAR_F32", "PVD_CNN_EPI_ACCUM_F32",
              "PVD_CNN_MAXIMA_FLOATS"]
    for c in consts:
        lines.append(f'printf("{c} %d\\n", (int)({c}));')
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = dict(line.split() for line in subprocess.run([str(exe)], capture_output=True, text=True,
                                                        check=True).stdout.splitlines())
    for name, (cls, fl) in fields.items():
        assert int(got[name]) == ctypes.sizeof(cls), name
        for f in fl:
            assert int(got[f"{name}.{f}"]) == getattr(cls, f).offset, (name, f)
    for c in consts:
        assert int(got[c]) == getattr(nat, c), c


def test_subregion_box_centres_on_the_maximum_and_stays_inside():
    import torch
    from pyvoxeldosimetry_amd.learned.pipeline import subregion_box
    a = torch.zeros((100, 80, 70))
    a[50, 40, 35] = 1.0
    assert subregion_box(a, 64) == (slice(18, 82), slice(8, 72), slice(3, 67))
    a.zero_()
    a[2, 79, 69] = 1.0                      # near faces: shifted inside
    assert subregion_box(a, 64) == (slice(0, 64), slice(16, 80), slice(6, 70))
    with pytest.raises(ValueError):
        subregion_box(torch.zeros((32, 80, 70)), 64)
