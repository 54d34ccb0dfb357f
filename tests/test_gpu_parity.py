"""GPU parity: the HIP path (through the C ABI) against the FP64 oracle.

Tolerance: the north-star bound for FP32 convolution, max|D - D64| <= 1e-5 *
max|D64| (``MAX_REL``); the dose-point kernels are FP64 on both sides
(relative 1e-12).
"""
import sys

import numpy as np
import pytest

from oracle import conv, dpk

pytestmark = pytest.mark.gpu

MAX_REL = 1e-5


def _say(*a):
    print("[gpu-case]", *a, file=sys.stderr, flush=True)


def dev_tensor(x):
    import torch
    return torch.from_numpy(np.ascontiguousarray(x, dtype=np.float32)).cuda()


def assert_parity(got, ref, tol=MAX_REL, what=""):
    got = np.asarray(got, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    assert got.shape == ref.shape, what
    assert np.isfinite(got).all(), f"{what}: non-finite output"
    m = conv.error_metrics(got, ref)
    assert m["max_rel"] <= tol, f"{what}: max_rel={m['max_rel']:.3e} rms_rel={m['rms_rel']:.3e}"
    return m


CONV_CASES = ["ref_small_aniso", "ref_crop", "ref_nonsmooth", "ref_prime", "ref_y90_24", "ref_lu177_even_kernel",
              "ref_density_scale", "lin_small", "lin_big_kernel", "lin_even_kernel", "lin_density", "ts_trapezoid",
              "ts_uneven", "ref_single_voxel", "lin_thin"]


def run_case(case, algo):
    from pyvoxeldosimetry_amd import ConvPlan
    acts = case["acts"]
    mode = str(case["mode"])
    plan = ConvPlan(acts.shape[1:], case["kernel"].shape, mode=mode, algo=algo)
    plan.set_kernel(case["kernel"])
    dens = dev_tensor(case["density"]) if "density" in case else None
    scale = float(case["scale"])
    if "times" in case:
        out = plan.dose_timeseries([dev_tensor(a) for a in acts], list(case["times"]), density=dens, scale=scale)
    else:
        out = plan.dose(dev_tensor(acts[0]), density=dens, scale=scale)
    res = out.cpu().numpy()
    plan.close()
    return res


@pytest.mark.parametrize("name", CONV_CASES)
@pytest.mark.parametrize("algo", ["fft", "direct"])
def test_golden_conv(golden, name, algo):
    case = golden[name]
    if algo == "direct":
        mode = str(case["mode"])
        shp = case["acts"].shape[1:]
        ku = [min(k, n) if mode == "reference" else k for k, n in zip(case["kernel"].shape, shp)]
        if max(ku) > 16:
            pytest.skip("kernel too large for the direct path")
    _say(name, algo)
    assert_parity(run_case(case, algo), case["expect"], what=f"{name}/{algo}")


def test_golden_nan_kernel_propagates(golden):
    _say("ref_y90_nan")
    case = golden["ref_y90_nan"]
    assert np.isnan(case["expect"]).all()
    for algo in ("fft", "direct"):
        got = run_case(case, algo)
        assert np.isnan(got).all(), algo


@pytest.mark.parametrize("name", ["dpk_y90_water_fix", "dpk_y90_bone_raw", "dpk_y90_lung_even", "dpk_lu177_water",
                                  "dpk_lu177_lung", "dpk_lu177_unknown_tissue", "dpk_ga68_water_fix",
                                  "dpk_ga68_lung_raw", "dpk_ga68_bone_even"])
def test_golden_dose_point_kernels(golden, name):
    from pyvoxeldosimetry_amd.dose_kernels import GENERATORS
    _say(name)
    c = golden[name]
    nuc = str(c["nuclide"])
    gen = GENERATORS[nuc](str(c["tissue"]), fix_centre=bool(c["fix_centre"])) if nuc in ("Y90", "Ga68") else \
        GENERATORS[nuc](str(c["tissue"]))
    got = gen.generate_kernel(float(c["voxel"]), tuple(int(g) for g in c["grid"]))
    exp = c["expect"]
    np.testing.assert_array_equal(np.isnan(got), np.isnan(exp))
    fin = np.isfinite(exp)
    np.testing.assert_allclose(got[fin], exp[fin], rtol=1e-12, atol=0)


# --- full-size configurations ----------------------------------------------------------
def test_small_config_64cubed_y90(capsys=None):
    """64^3 at 4.42 mm, Y-90 11^3 (fixed centre), decay-scaled, reference and linear modes."""
    from pyvoxeldosimetry_amd import ConvPlan, Y90KernelGenerator
    from pyvoxeldosimetry_amd.phantoms import small_phantom
    _say("small config 64^3")
    p = small_phantom()
    k = Y90KernelGenerator("water", fix_centre=True).generate_kernel_device(p.voxel_mm, p.kernel_grid)
    knp = k.cpu().numpy()
    decay = 2.0 ** (-2.0 / 64.1)
    a = dev_tensor(p.activity[0])
    for mode, oracle_fn in (("reference", conv.ref_conv), ("linear", conv.linear_same)):
        exp = oracle_fn(p.activity[0].astype(np.float64), knp) * decay
        for algo in ("fft", "direct"):
            plan = ConvPlan(p.shape, p.kernel_grid, mode=mode, algo=algo)
            plan.set_kernel(k)
            got = plan.dose(a, scale=decay).cpu().numpy()
            m = assert_parity(got, exp, what=f"small/{mode}/{algo}")
            _say(f"small {mode} {algo} max_rel={m['max_rel']:.2e} rms_rel={m['rms_rel']:.2e}")
            plan.close()


def test_ct_config_256x256x200_lu177_density():
    """256x256x200 at 2 mm, Lu-177 31^3, density scaling; FP64 oracle of the reference expression."""
    from pyvoxeldosimetry_amd import ConvPlan, Lu177KernelGenerator
    from pyvoxeldosimetry_amd.phantoms import ct_phantom
    _say("ct config 256x256x200")
    p = ct_phantom()
    k = Lu177KernelGenerator("water").generate_kernel_device(p.voxel_mm, p.kernel_grid)
    exp = conv.epilogue(conv.ref_conv(p.activity[0].astype(np.float64), k.cpu().numpy()), p.density)
    plan = ConvPlan(p.shape, p.kernel_grid, mode="reference", algo="fft")
    plan.set_kernel(k)
    got = plan.dose(dev_tensor(p.activity[0]), density=dev_tensor(p.density)).cpu().numpy()
    m = assert_parity(got, exp, what="ct/reference/fft")
    _say(f"ct reference max_rel={m['max_rel']:.2e} rms_rel={m['rms_rel']:.2e} fft_n={plan.fft_n}")


def test_ct_config_linear_mode():
    from pyvoxeldosimetry_amd import ConvPlan, Lu177KernelGenerator
    from pyvoxeldosimetry_amd.phantoms import ct_phantom
    _say("ct config linear")
    p = ct_phantom()
    k = Lu177KernelGenerator("water").generate_kernel_device(p.voxel_mm, p.kernel_grid)
    exp = conv.epilogue(conv.linear_same(p.activity[0].astype(np.float64), k.cpu().numpy()), p.density)
    plan = ConvPlan(p.shape, p.kernel_grid, mode="linear", algo="fft")
    plan.set_kernel(k)
    got = plan.dose(dev_tensor(p.activity[0]), density=dev_tensor(p.density)).cpu().numpy()
    m = assert_parity(got, exp, what="ct/linear/fft")
    _say(f"ct linear max_rel={m['max_rel']:.2e} fft_n={plan.fft_n}")


@pytest.mark.parametrize("mode", ["reference", "linear"])
def test_symmetric_and_general_spectra_agree(mode):
    """Real-octant spectrum (odd symmetric kernel) vs the full complex spectrum vs the oracle,
    on a smooth size (reference: P = N, output roll) and a prime size (padded + fold)."""
    from pyvoxeldosimetry_amd import ConvPlan, Lu177KernelGenerator
    _say("symmetric spectrum", mode)
    rng = np.random.default_rng(21)
    k = Lu177KernelGenerator("water").generate_kernel_device(1.0, (9, 9, 9))
    for shape in ((40, 36, 30), (29, 31, 23)):
        a = rng.random(shape).astype(np.float32)
        oracle_fn = conv.ref_conv if mode == "reference" else conv.linear_same
        exp = oracle_fn(a.astype(np.float64), k.cpu().numpy())
        outs = []
        for general in (False, True):
            plan = ConvPlan(shape, (9, 9, 9), mode=mode, algo="fft", general_spectrum=general)
            plan.set_kernel(k)
            assert plan.symmetric_spectrum == (not general)
            outs.append(plan.dose(dev_tensor(a)).cpu().numpy())
            assert_parity(outs[-1], exp, what=f"{mode} {shape} general={general}")
    # non-symmetric kernels keep the general spectrum
    plan = ConvPlan((20, 20, 20), (5, 5, 5), algo="fft")
    plan.set_kernel(rng.random((5, 5, 5)))
    assert not plan.symmetric_spectrum


def test_nonsmooth_reference_fold_against_bruteforce_c():
    """Prime-sized grid: padded FFT + fold must equal circular brute force (independent C)."""
    from oracle import brute
    from pyvoxeldosimetry_amd import ConvPlan
    _say("fold vs C brute force 29x31x37")
    rng = np.random.default_rng(11)
    a = rng.random((29, 31, 37)).astype(np.float32)
    k = rng.random((9, 9, 9))
    exp = brute.brute_conv3d(a.astype(np.float64), k, "reference")
    for fft_n in (None, (40, 40, 48)):
        plan = ConvPlan(a.shape, k.shape, mode="reference", algo="fft", fft_n=fft_n)
        plan.set_kernel(k)
        assert_parity(plan.dose(dev_tensor(a)).cpu().numpy(), exp, what=f"fold {fft_n}")


# --- properties at full size -------------------------------------------------------------
def test_delta_returns_kernel_at_full_size():
    import torch
    from pyvoxeldosimetry_amd import ConvPlan, Lu177KernelGenerator
    _say("delta 256x256x200")
    shape = (256, 256, 200)
    k = Lu177KernelGenerator("water").generate_kernel_device(2.0, (31, 31, 31))
    plan = ConvPlan(shape, (31, 31, 31), mode="reference", algo="fft")
    plan.set_kernel(k)
    a = torch.zeros(shape, dtype=torch.float32, device="cuda")
    a[100, 17, 190] = 1.0
    got = plan.dose(a).cpu().numpy()
    exp = np.roll(conv.pad_or_crop(k.cpu().numpy(), shape), (100, 17, 190), axis=(0, 1, 2))
    assert_parity(got, exp, what="delta")


def test_linearity_at_full_size():
    import torch
    from pyvoxeldosimetry_amd import ConvPlan, Lu177KernelGenerator
    _say("linearity 256x256x200")
    shape = (256, 256, 200)
    k = Lu177KernelGenerator("water").generate_kernel_device(2.0, (31, 31, 31))
    plan = ConvPlan(shape, (31, 31, 31), mode="reference", algo="fft")
    plan.set_kernel(k)
    g = torch.Generator(device="cuda").manual_seed(3)
    a = torch.rand(shape, device="cuda", generator=g)
    b = torch.rand(shape, device="cuda", generator=g)
    da, db = plan.dose(a), plan.dose(b)
    dab = plan.dose_weighted([a, b], [2.0, -0.5])
    ref = 2.0 * da.double() - 0.5 * db.double()
    err = (dab.double() - ref).abs().max().item() / ref.abs().max().item()
    assert err < MAX_REL
    # constant input -> constant * sum(kernel)
    c = plan.dose(torch.full(shape, 3.0, device="cuda")).double()
    exp = 3.0 * k.sum().item()
    assert (c - exp).abs().max().item() / exp < MAX_REL


def test_repeat_calls_are_deterministic_and_out_param():
    import torch
    from pyvoxeldosimetry_amd import ConvPlan
    _say("determinism")
    rng = np.random.default_rng(5)
    a = dev_tensor(rng.random((40, 36, 30)))
    plan = ConvPlan(a.shape, (7, 7, 7), algo="fft")
    plan.set_kernel(rng.random((7, 7, 7)))
    o1 = plan.dose(a)
    o2 = torch.empty_like(a)
    plan.dose(a, out=o2)
    assert torch.equal(o1, o2)


def test_weighted_sum_kernel():
    import torch
    from pyvoxeldosimetry_amd.core.activity_sampler import weighted_sum_device
    _say("weighted sum")
    rng = np.random.default_rng(2)
    maps = [rng.random((17, 9, 13)).astype(np.float32) for _ in range(5)]
    w = [0.5, -1.25, 3.0, 0.0, 7.5]
    got = weighted_sum_device([dev_tensor(m) for m in maps], w).cpu().numpy()
    exp = sum(wi * m.astype(np.float64) for wi, m in zip(w, maps))
    np.testing.assert_allclose(got, exp, rtol=1e-6, atol=1e-5)
    # unaligned tail path
    base = torch.from_numpy(np.arange(1001, dtype=np.float32)).cuda()
    got = weighted_sum_device([base[1:], base[1:]], [1.0, 2.0]).cpu().numpy()
    np.testing.assert_allclose(got, 3.0 * np.arange(1, 1001), rtol=1e-7)


def test_plan_rejects_bad_inputs():
    import torch
    from pyvoxeldosimetry_amd import ConvPlan
    _say("plan validation")
    with pytest.raises(ValueError):
        ConvPlan((8, 8, 8), (3, 3, 3), fft_n=(9, 8, 8))        # aliases in reference mode
    with pytest.raises(ValueError):
        ConvPlan((8, 8, 8), (3, 3, 3), fft_n=(11, 8, 8))       # not 7-smooth
    with pytest.raises(NotImplementedError):
        ConvPlan((40, 8, 8), (31, 3, 3), mode="linear", algo="direct")     # 31 taps used: > 16
    # a 31-wide kernel on an 8-wide grid only ever uses 15 taps in linear mode
    assert ConvPlan((8, 8, 8), (31, 3, 3), mode="linear", algo="direct").kernel_used == (15, 3, 3)
    plan = ConvPlan((8, 8, 8), (3, 3, 3))
    a = torch.zeros((8, 8, 8), device="cuda")
    with pytest.raises(ValueError, match="no kernel"):
        plan.dose(a)
    plan.set_kernel(np.ones((3, 3, 3)))
    with pytest.raises(ValueError):
        plan.dose(torch.zeros((8, 8, 9), device="cuda"))
    with pytest.raises(ValueError):
        plan.dose(a.double())


def test_torso_config_512x512x400_61cubed_6_timepoints():
    """Full-resolution volume: linear (padded) FFT, Lu-177 61^3, T=6, density; FP64 oracle."""
    from pyvoxeldosimetry_amd import ConvPlan, Lu177KernelGenerator
    from pyvoxeldosimetry_amd.phantoms import torso_phantom
    _say("torso 512x512x400 T=6")
    ph = torso_phantom()
    k = Lu177KernelGenerator("water").generate_kernel_device(ph.voxel_mm, ph.kernel_grid)
    plan = ConvPlan(ph.shape, ph.kernel_grid, mode="linear", algo="fft")
    plan.set_kernel(k)
    got = plan.dose_timeseries([dev_tensor(a) for a in ph.activity], ph.time_points_h,
                               density=dev_tensor(ph.density)).cpu().numpy()
    w = conv.trapezoid_weights_seconds(ph.time_points_h)
    acc = sum(wi * a.astype(np.float64) for wi, a in zip(w, ph.activity))
    exp = conv.epilogue(conv.linear_same(acc, k.cpu().numpy()), ph.density)
    m = assert_parity(got, exp, what="torso/linear")
    _say(f"torso max_rel={m['max_rel']:.2e} rms_rel={m['rms_rel']:.2e} fft_n={plan.fft_n}")


@pytest.mark.parametrize("mode,shape,kshape,sym", [
    ("linear", (241, 240, 40), (31, 31, 31), True),       # P = 256 on x and y, off = 0, zero rows past N
    ("reference", (256, 256, 24), (9, 7, 5), False),     # P = N = 256, complex spectrum, origin anchor
    ("reference", (256, 100, 20), (11, 11, 11), True),    # 256 on x only (y takes the generic path)
])
def test_register_256_point_path(mode, shape, kshape, sym):
    """The register-resident 256-point y/x transforms: linear padding, a non-symmetric
    kernel (general complex spectrum) and a mix with the generic path, vs the oracle."""
    from pyvoxeldosimetry_amd import ConvPlan
    _say("256-point register path", mode, shape, kshape)
    rng = np.random.default_rng(sum(shape))
    a = rng.random(shape).astype(np.float32)
    rho = rng.uniform(0.5, 1.5, size=shape).astype(np.float32)
    k = dpk.lu177_kernel(1.0, kshape) if sym else rng.random(kshape)
    plan = ConvPlan(shape, kshape, mode=mode, algo="fft")
    plan.set_kernel(k)
    assert 256 in plan.fft_n[:2] and plan.symmetric_spectrum == sym
    oracle_fn = conv.ref_conv if mode == "reference" else conv.linear_same
    exp = conv.epilogue(oracle_fn(a.astype(np.float64), np.asarray(k, np.float64)), rho)
    got = plan.dose(dev_tensor(a), density=dev_tensor(rho)).cpu().numpy()
    assert_parity(got, exp, what=f"{mode} {shape} {kshape}")


@pytest.mark.parametrize("tile,max_radix", [((8, 8, 8), 16), ((32, 32, 32), 8), ((16, 32, 8), 8)])
def test_tile_and_radix_choices_keep_parity(tile, max_radix):
    """Every tile shape / radix the autotuner may pick gives the oracle's result
    (256-point axes fall back to the generic path when their tile is not 16)."""
    from pyvoxeldosimetry_amd import ConvPlan
    _say("tiles/radix", tile, max_radix)
    rng = np.random.default_rng(5)
    k = dpk.lu177_kernel(2.0, (9, 9, 9))
    for shape in ((256, 64, 40), (45, 52, 33)):
        a = rng.random(shape).astype(np.float32)
        plan = ConvPlan(shape, (9, 9, 9), mode="reference", algo="fft", tile=tile, max_radix=max_radix)
        plan.set_kernel(k)
        info = plan.info()
        assert info["max_radix"] == max_radix and all(t <= w for t, w in zip(info["tile"], tile))
        assert_parity(plan.dose(dev_tensor(a)).cpu().numpy(), conv.ref_conv(a.astype(np.float64), k),
                      what=f"tiles {tile} radix {max_radix} {shape}")


# z lengths with a register two-step transform P = A * B (A = 8 or 16), each in
# reference mode (P = Nz) and one in linear mode (P chosen by the plan)
ZREG_LENGTHS = [96, 112, 120, 128, 144, 160, 168, 192, 200, 216, 224, 240, 256, 288, 320, 336, 384, 400, 432, 480]


@pytest.mark.parametrize("nz", ZREG_LENGTHS)
def test_register_z_two_step_against_oracle_and_stockham(nz):
    """Every z length with a register two-step transform matches the oracle, with a
    symmetric (real-octant) and a general kernel, and gives the LDS Stockham
    stages' result within FP32 rounding."""
    from pyvoxeldosimetry_amd import ConvPlan
    _say("register z two-step", nz)
    rng = np.random.default_rng(nz)
    shape = (6, 5, nz)
    a = rng.random(shape).astype(np.float32)
    rho = rng.uniform(0.5, 1.5, size=shape).astype(np.float32)
    for kshape, sym in (((5, 5, 7), True), ((3, 4, 9), False)):
        k = dpk.lu177_kernel(1.0, kshape) if sym else rng.random(kshape)
        exp = conv.epilogue(conv.ref_conv(a.astype(np.float64), np.asarray(k, np.float64)), rho)
        outs = []
        for lds_only in (False, True):
            plan = ConvPlan(shape, kshape, mode="reference", algo="fft", stockham=lds_only)
            plan.set_kernel(k)
            assert plan.fft_n[2] == nz
            got = plan.dose(dev_tensor(a), density=dev_tensor(rho)).cpu().numpy()
            plan.close()
            assert_parity(got, exp, what=f"nz={nz} k={kshape} stockham={lds_only}")
            outs.append(got)
        assert np.abs(outs[0] - outs[1]).max() <= 1e-6 * np.abs(exp).max()


@pytest.mark.parametrize("smooth_lengths", [False, True])
def test_register_z_two_step_linear_and_batch(smooth_lengths):
    """Linear mode picks padded z lengths: 160 and 200 are register lengths; for 200 + 7
    the 7-smooth rule gives 210 (LDS Stockham) and the default pads to 216 = 8 x 27 (a
    register length).  A batch of volumes with different z lengths runs each with its
    own transform."""
    from pyvoxeldosimetry_amd import ConvPlan
    from pyvoxeldosimetry_amd.batch import DoseBatch
    rng = np.random.default_rng(77)
    k = dpk.lu177_kernel(2.0, (15, 15, 15))
    shapes = [(20, 18, 146), (16, 24, 190), (9, 11, 200)]      # linear "same": P >= Nz + 7
    expect_z = [160, 200, 210 if smooth_lengths else 216]
    acts = [rng.random(s).astype(np.float32) for s in shapes]
    for s, a, pz in zip(shapes, acts, expect_z):
        _say("register z linear", s, smooth_lengths)
        plan = ConvPlan(s, k.shape, mode="linear", algo="fft", smooth_lengths=smooth_lengths)
        plan.set_kernel(k)
        assert plan.fft_n[2] == pz, plan.fft_n
        assert (plan.fft_path[2] == "two_step") == (pz != 210)
        assert_parity(plan.dose(dev_tensor(a)).cpu().numpy(), conv.linear_same(a.astype(np.float64), k),
                      what=f"linear {s} fft_n={plan.fft_n}")
        plan.close()
    _say("register z batch")
    b = DoseBatch(shapes, [k] * len(shapes), mode="linear", smooth_lengths=smooth_lengths)
    for i, a in enumerate(acts):
        b.set_volume(i, [dev_tensor(a)], [1.0])
    outs = b.run(graph=True)
    for s, a, o in zip(shapes, acts, outs):
        assert_parity(o.cpu().numpy(), conv.linear_same(a.astype(np.float64), k), what=f"batch {s}")


# y/x lengths P = 16 * B with a register two-step transform
XY_TWO_STEP_B = [6, 8, 9, 10, 12, 14, 15, 18, 20, 21, 24, 25, 27, 28, 30, 32, 35, 36]


@pytest.mark.parametrize("b", XY_TWO_STEP_B)
def test_xy_register_two_step_against_oracle_and_stockham(b):
    """The y and x passes of length P = 16 * B (not 256) take the register two-step
    (natural to permuted order and back, with the permuted stores): reference mode
    with P = N on either axis, and padded reference mode (fft_n > N: circular fold
    and origin offset on the output rows), symmetric and general kernels; each
    matches the FP64 oracle and the LDS Stockham stages within FP32 rounding."""
    from pyvoxeldosimetry_amd import ConvPlan
    P = 16 * b
    rng = np.random.default_rng(b)
    cases = [((P, 6, 10), (5, 3, 3), None, True), ((7, P, 10), (3, 5, 3), None, False),
             ((P - 9, 5, 8), (7, 3, 3), (P, 0, 0), False), ((5, P - 5, 8), (3, 5, 3), (0, P, 0), True)]
    for shape, kshape, fft_n, sym in cases:
        _say("xy two-step", P, shape, kshape, fft_n)
        a = rng.random(shape).astype(np.float32)
        rho = rng.uniform(0.5, 1.5, size=shape).astype(np.float32)
        k = dpk.lu177_kernel(1.0, kshape) if sym else rng.random(kshape)
        exp = conv.epilogue(conv.ref_conv(a.astype(np.float64), np.asarray(k, np.float64)), rho)
        outs = []
        for lds_only in (False, True):
            plan = ConvPlan(shape, kshape, mode="reference", algo="fft", fft_n=fft_n, stockham=lds_only)
            plan.set_kernel(k)
            assert P in plan.fft_n[:2], plan.fft_n
            got = plan.dose(dev_tensor(a), density=dev_tensor(rho)).cpu().numpy()
            plan.close()
            assert_parity(got, exp, what=f"P={P} {shape} fft_n={fft_n} stockham={lds_only}")
            outs.append(got)
        assert np.abs(outs[0] - outs[1]).max() <= 1e-6 * np.abs(exp).max()


def _offset_copy(a, off):
    """A contiguous float32 device copy of `a` starting `off` floats into its storage."""
    import torch
    buf = torch.empty(a.size + off, dtype=torch.float32, device="cuda")
    view = buf[off:].view(a.shape)
    view.copy_(torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)))
    return view


@pytest.mark.parametrize("shape,tile", [((9, 7, 13), None), ((20, 11, 31), None), ((6, 5, 40), None),
                                        ((9, 7, 13), (16, 16, 1))])
def test_fz_load_forms_agree_bit_for_bit(shape, tile):
    """FZ reads the T activity maps as float4 along each line (Nz % 4 == 0, 16-byte
    aligned maps), as float4 runs of a tile's contiguous lines (aligned maps, any Nz,
    tiles of >= 2 line pairs) or as 4-byte loads (maps not 16-byte aligned, or 1-pair
    tiles).  The same maps, 16-byte aligned and one float off, give the FP64 oracle's
    weighted dose and the same bits."""
    import torch
    from pyvoxeldosimetry_amd import ConvPlan
    _say("fz load forms", shape, tile)
    rng = np.random.default_rng(sum(shape))
    acts = [rng.random(shape).astype(np.float32) for _ in range(3)]
    w = [0.5, 2.0, 1.5]
    k = dpk.lu177_kernel(2.0, (5, 7, 9))
    plan = ConvPlan(shape, k.shape, mode="reference", algo="fft", tile=tile)
    plan.set_kernel(k)
    exp = conv.ref_conv(sum(wi * a.astype(np.float64) for wi, a in zip(w, acts)), k)
    outs = []
    for off in (0, 1):
        maps = [_offset_copy(a, off) if off else dev_tensor(a) for a in acts]
        assert all((m.data_ptr() % 16 == 0) == (off == 0) for m in maps)
        got = plan.dose_weighted(maps, w).cpu().numpy()
        assert_parity(got, exp, what=f"fz load {shape} tile {tile} offset {off}")
        outs.append(got)
    torch.cuda.synchronize()
    plan.close()
    assert np.array_equal(outs[0], outs[1])


def test_xy_two_step_linear_time_series_with_density():
    """Linear mode (no fold, offset 0) on a 16 * B length (y: 313 + 7 -> 320 = 16 x 20)
    with a 3-timepoint trapezoid and density: the time-integrated dose matches the FP64
    oracle."""
    from pyvoxeldosimetry_amd import ConvPlan
    _say("xy two-step linear time series")
    rng = np.random.default_rng(320)
    shape, kshape = (29, 313, 24), (15, 15, 5)
    acts = [rng.random(shape).astype(np.float32) for _ in range(3)]
    rho = rng.uniform(0.5, 1.5, size=shape).astype(np.float32)
    t_h = [1.0, 4.0, 24.0]
    k = dpk.lu177_kernel(2.0, kshape)
    plan = ConvPlan(shape, kshape, mode="linear", algo="fft")
    plan.set_kernel(k)
    assert plan.fft_n[1] == 320, plan.fft_n
    got = plan.dose_timeseries([dev_tensor(a) for a in acts], t_h, density=dev_tensor(rho)).cpu().numpy()
    plan.close()
    w = conv.trapezoid_weights_seconds(t_h)
    acc = sum(wi * a.astype(np.float64) for wi, a in zip(w, acts))
    assert_parity(got, conv.epilogue(conv.linear_same(acc, k), rho), what="linear two-step time series")


@pytest.mark.parametrize("shape,kshape", [((1, 320, 2), (1, 5, 1)), ((2, 3, 400), (1, 3, 9)), ((400, 1, 3), (7, 1, 1))])
def test_register_paths_on_thin_volumes(shape, kshape):
    """Register two-step lengths next to 1-, 2- and 3-voxel axes (reference mode, P = N):
    the dose matches the FP64 oracle."""
    from pyvoxeldosimetry_amd import ConvPlan
    _say("thin volume", shape, kshape)
    rng = np.random.default_rng(sum(shape))
    a = rng.random(shape).astype(np.float32)
    k = rng.random(kshape)
    plan = ConvPlan(shape, kshape, mode="reference", algo="fft")
    plan.set_kernel(k)
    got = plan.dose(dev_tensor(a)).cpu().numpy()
    plan.close()
    assert_parity(got, conv.ref_conv(a.astype(np.float64), k), what=f"thin {shape}")
