"""GPU tests at the sizes of the benchmark configurations that no other test runs.

Batched uneven workload (48 seeded phantoms, sizes U{96..384}^2 x U{96..320}, T = 4,
mixed Y-90 11^3 / Lu-177 31^3, reference convolution): every volume of the batch
bit-exact against its own single-volume plan, in eager launches and in the graph,
and six named members (both nuclides, the largest volume) against the FP64
oracle of kernel_convolution.py:78-106 (the trapezoid over per-timepoint dose
rates, taken by linearity as one convolution of the trapezoid-weighted activity).

Learned correction at 128^3: the residual CNN's forward on the GPU against the
FP64 network oracle on windows (six 3^3 layers see 6 voxels, so a 20^3 input
window pins its central 8^3 outputs, and a window at the volume edge also the
outputs next to the edge), then the corrected dose with its fallback rule, and the
network run in z slabs equal to the whole-volume network.

Full-resolution torso (512x512x400 at 0.98 mm, 6 timepoints, Lu-177): the
time-integrated dose corrected by the Lu-177 network in z slabs (400 is not a
multiple of 32), with its bounded-correction rule, and against Monte Carlo on a 64^3
sub-region (activity kept only inside the box, the plain dose and the correction run
on the whole volume).
"""
import sys

import numpy as np
import pytest

from oracle import cnn as ocnn
from oracle import conv

pytestmark = pytest.mark.gpu

MAX_REL = 1e-5


def _say(*a):
    print("[gpu-case]", *a, file=sys.stderr, flush=True)


# --------------------------------------------------------------------- batched workload
CHECKED = [23, 5, 0, 4, 14, 37]     # largest (376x276x315 Y-90), largest Lu-177, and four more


@pytest.fixture(scope="module")
def batch48():
    import torch
    from pyvoxeldosimetry_amd import Lu177KernelGenerator, Y90KernelGenerator
    from pyvoxeldosimetry_amd.batch import DoseBatch
    from pyvoxeldosimetry_amd.phantoms import BATCH_TIMEPOINTS_H, batch_phantom, batch_specs
    specs = batch_specs(seed=3)
    assert len(specs) == 48
    kern = {"Y90": Y90KernelGenerator("water", fix_centre=True).generate_kernel_device(2.0, (11, 11, 11)),
            "Lu177": Lu177KernelGenerator("water").generate_kernel_device(2.0, (31, 31, 31))}
    _say("batch48: plans")
    b = DoseBatch([s.shape for s in specs], [kern[s.nuclide] for s in specs], mode="reference")
    acts = []
    for i, s in enumerate(specs):
        ph = batch_phantom(s)
        acts.append([torch.from_numpy(a).cuda() for a in ph.activity])
        b.set_volume_timeseries(i, acts[i], BATCH_TIMEPOINTS_H)
        if i % 12 == 11:
            _say(f"batch48: {i + 1} volumes uploaded")
    yield specs, kern, b, acts
    b.close()


def test_batch48_every_volume_bit_exact_with_its_single_plan(batch48):
    import torch
    from pyvoxeldosimetry_amd import ConvPlan
    from pyvoxeldosimetry_amd.phantoms import BATCH_TIMEPOINTS_H
    specs, kern, b, acts = batch48
    eager = [o.clone() for o in b.run(graph=False)]
    graph = [o.clone() for o in b.run(graph=True)]
    torch.cuda.synchronize()
    for i, s in enumerate(specs):
        k = kern[s.nuclide]
        plan = ConvPlan(s.shape, tuple(k.shape), mode="reference", algo="fft")
        plan.set_kernel(k)
        single = plan.dose_timeseries(acts[i], BATCH_TIMEPOINTS_H)
        torch.cuda.synchronize()
        assert torch.equal(eager[i], single), f"volume {i} {s.shape} eager, fft_n {plan.fft_n}"
        assert torch.equal(graph[i], single), f"volume {i} {s.shape} graph, fft_n {plan.fft_n}"
        plan.close()
        if i % 12 == 11:
            _say(f"batch48: {i + 1} volumes bit-exact")


def test_batch48_named_members_against_the_fp64_oracle(batch48):
    import torch
    from pyvoxeldosimetry_amd.phantoms import BATCH_TIMEPOINTS_H, batch_phantom
    specs, kern, b, _ = batch48
    outs = b.run(graph=True)
    torch.cuda.synchronize()
    nuclides = {specs[i].nuclide for i in CHECKED}
    assert nuclides == {"Y90", "Lu177"}
    assert max(CHECKED, key=lambda i: np.prod(specs[i].shape)) == max(range(48), key=lambda i: np.prod(specs[i].shape))
    w = conv.trapezoid_weights_seconds(BATCH_TIMEPOINTS_H)
    for i in CHECKED:
        s = specs[i]
        ph = batch_phantom(s)
        acc = sum(wi * a.astype(np.float64) for wi, a in zip(w, ph.activity))
        exp = conv.ref_conv_threaded(acc, kern[s.nuclide].cpu().numpy(), workers=16)
        m = conv.error_metrics(outs[i].cpu().numpy(), exp)
        _say(f"batch48 member {i} {s.shape} {s.nuclide}: max_rel {m['max_rel']:.2e} rms_rel {m['rms_rel']:.2e}")
        assert m["max_rel"] <= MAX_REL, (i, s.shape, m)


# --------------------------------------------------------------------- learned correction, 128^3
@pytest.fixture(scope="module")
def cnn128():
    from pyvoxeldosimetry_amd.learned.pipeline import LearnedCorrector, synth_phantom
    corr = LearnedCorrector.pretrained("Y90", 2.0)
    act, rho = synth_phantom(20001, (128, 128, 128), corr.device)
    return corr, act, rho


def _valid(start, n=128, w=20, rf=6):
    """Outputs of a window whose 6-voxel neighbourhood lies in the window or outside the volume."""
    lo = 0 if start == 0 else rf
    hi = w if start + w == n else w - rf
    return lo, hi


@pytest.mark.parametrize("start", [(54, 60, 57), (0, 0, 0), (108, 108, 108), (0, 64, 108)])
def test_cnn_forward_128cubed_matches_the_network_oracle_on_windows(cnn128, start):
    corr, act, rho = cnn128
    shape = (128, 128, 128)
    _say(f"cnn 128^3 window {start}")
    dp = corr.plain(act, rho)
    x, _ = corr.pack(act, rho, dp)
    r = corr.net.forward(x, shape, keep_activations=False).reshape(shape).cpu().numpy().astype(np.float64)
    xh = x.float().cpu().numpy().astype(np.float64).reshape(shape + (16,))
    params = corr.net.params_numpy()
    sl = tuple(slice(s, s + 20) for s in start)
    ro = ocnn.forward_stored(params, xh[sl])
    v = [_valid(s) for s in start]
    got = r[tuple(slice(s + lo, s + hi) for s, (lo, hi) in zip(start, v))]
    exp = ro[tuple(slice(lo, hi) for lo, hi in v)]
    scale = max(np.abs(exp).max(), 1e-6)
    diff = np.abs(got - exp).max() / scale
    fro = np.linalg.norm(got - exp) / max(np.linalg.norm(exp), 1e-30)
    _say(f"window {start}: |r| max {scale:.3e}, max diff {diff:.2e} of max, rel fro {fro:.2e}")
    assert np.abs(exp).max() > 1e-4                    # the trained network corrects here
    assert diff <= 2e-2 and fro <= 5e-3


@pytest.mark.parametrize("region", [True, False])
def test_corrected_dose_128cubed_applies_the_bounded_correction(cnn128, region):
    import torch
    corr, act, rho = cnn128
    _say("cnn 128^3 corrected dose", region)
    c = corr.correct(act, rho, region=region)
    assert not c.fallback
    applied = c.plain >= (corr.region_frac if region else 0.0) * c.plain.max()
    assert torch.isfinite(c.r[applied]).all() and torch.isfinite(c.dose).all()
    ok = (c.r.abs() <= corr.bound) & applied
    expected = torch.where(ok, c.plain * (1.0 + c.r), c.plain)
    torch.testing.assert_close(c.dose, expected, rtol=2e-6, atol=0)
    ten = c.plain >= 0.1 * c.plain.max()
    assert c.exceed == int((~(c.r.abs() <= corr.bound) & ten).sum().item())


@pytest.mark.parametrize("seed", [20000, 20002])
def test_region_restricted_network_equals_the_full_network_on_its_region(seed):
    """The network run only on the blocks within its 6-voxel receptive field of the
    region dp >= 5 % of max gives the full-volume network's output bit for bit there;
    outside, the dose is the plain dose.  Seed 20000 has a small region (~1 % of the
    volume), 20002 a large one (~94 %)."""
    import torch
    from pyvoxeldosimetry_amd.learned.pipeline import LearnedCorrector, synth_phantom
    _say("cnn region vs full", seed)
    corr = LearnedCorrector.pretrained("Y90", 2.0)
    act, rho = synth_phantom(seed, (128, 128, 128), corr.device)
    full = corr.correct(act, rho, region=False)
    full_dose, full_r = full.dose.clone(), full.r.clone()
    part = corr.correct(act, rho, region=True)
    reg = part.plain >= corr.region_frac * part.plain.max()
    assert reg.any()
    assert torch.equal(part.r[reg], full_r[reg])
    assert torch.equal(part.dose[reg], full_dose[reg])
    assert torch.equal(part.dose[~reg], part.plain[~reg])


def test_tiled_network_equals_the_whole_volume_network(cnn128):
    """The network run in z slabs (extent 64, kept outputs >= 6 voxels from cut faces,
    inputs normalised by the whole volume) gives the whole-volume network's output bit
    for bit; on a 100-voxel-deep volume (not a multiple of 32, so only the slab path
    runs) it equals the whole network on the 96-deep crop wherever the crop's own cut
    face is more than 6 voxels away."""
    import torch
    corr, act, rho = cnn128
    shape = (128, 128, 128)
    _say("cnn tiled vs whole, 128^3")
    dp = corr.plain(act, rho)
    x, _ = corr.pack(act, rho, dp)
    full = corr.net.forward(x, shape, keep_activations=False).clone().reshape(shape)
    tiled, _ = corr.network_tiled(act, rho, dp, tile_z=64)
    assert torch.equal(tiled.reshape(shape), full)
    _say("cnn tiled, 128x128x100")
    a2, r2 = act[:, :, :100].contiguous(), rho[:, :, :100].contiguous()
    dp2 = corr.plain(a2, r2)
    t2, _ = corr.network_tiled(a2, r2, dp2, tile_z=64)
    t2 = t2.reshape(128, 128, 100)
    assert torch.isfinite(t2).all()
    # the whole network on the first 96 planes, normalised like the 100-deep volume
    a3, r3, d3 = a2[:, :, :96].contiguous(), r2[:, :, :96].contiguous(), dp2[:, :, :96].contiguous()
    x3, _ = corr.pack(a3, r3, d3, scales=(a2.max().reshape(1), dp2.max().reshape(1)))
    f3 = corr.net.forward(x3, (128, 128, 96), keep_activations=False).reshape(128, 128, 96)
    assert torch.equal(t2[:, :, :90], f3[:, :, :90])
    c = corr.correct(a2, r2)
    assert torch.isfinite(c.dose).all() and torch.equal(c.r, t2)


# --------------------------------------------------------------------- torso, tiled correction
@pytest.fixture(scope="module")
def torso_lu177():
    import torch
    from pyvoxeldosimetry_amd.core.activity_sampler import trapezoid_weights
    from pyvoxeldosimetry_amd.learned.pipeline import LearnedCorrector
    from pyvoxeldosimetry_amd.phantoms import torso_phantom
    _say("torso 512x512x400: phantom and Lu-177 corrector")
    ph = torso_phantom()
    corr = LearnedCorrector.pretrained("Lu177", ph.voxel_mm)
    cum = torch.zeros(ph.shape, dtype=torch.float32, device=corr.device)
    for w, a in zip(trapezoid_weights(ph.time_points_h), ph.activity):
        cum.add_(torch.from_numpy(a).to(corr.device), alpha=w * 3600.0)
    rho = torch.from_numpy(ph.density).to(corr.device)
    return corr, cum, rho


def test_torso_tiled_correction_applies_the_bounded_correction(torso_lu177):
    import torch
    corr, cum, rho = torso_lu177
    _say("torso tiled correction")
    c = corr.correct(cum, rho)
    assert not c.fallback
    assert torch.isfinite(c.r).all() and torch.isfinite(c.dose).all()
    # outside the body the plain dose is FFT round-off around 0: the correction applies
    # where it is >= 0 (the region threshold 0 * max)
    ok = (c.r.abs() <= corr.bound) & (c.plain >= 0)
    expected = torch.where(ok, c.plain * (1.0 + c.r), c.plain)
    torch.testing.assert_close(c.dose, expected, rtol=2e-6, atol=0)
    ten = c.plain >= 0.1 * c.plain.max()
    assert c.exceed == int((~(c.r.abs() <= corr.bound) & ten).sum().item())
    # the automatic slab extent (two slabs of 224 here) and slabs of 128: the same bits
    dose = c.dose.clone()
    assert torch.equal(corr.correct(cum, rho, tile_z=128).dose, dose)


def _torso_boxes(cum):
    """The 64^3 sub-regions: around the activity maximum (a lesion in soft tissue), and
    across the lower boundary of the right lung (lung 0.26 / soft tissue 1.04)."""
    from pyvoxeldosimetry_amd.learned.pipeline import subregion_box
    sx, sy, sz = cum.shape
    return {"hottest": subregion_box(cum, 64),
            "lung_edge": subregion_box(cum, 64, centre=(sx // 2, int(0.32 * sy), int(0.5 * sz)))}


def test_torso_correction_against_monte_carlo_on_64cubed_subregions(torso_lu177):
    """Soft-tissue lesion: the plain dose is already within ~1 % of Monte Carlo and the
    corrected dose stays within 2 % on average; lung boundary: the plain dose is far off
    (rho_w / rho scaling of a water kernel) and the correction removes most of it."""
    from pyvoxeldosimetry_amd.learned.pipeline import evaluate_subregion
    corr, cum, rho = torso_lu177
    res = {}
    for name, box in _torso_boxes(cum).items():
        _say("torso sub-region Monte Carlo", name, [(b.start, b.stop) for b in box])
        ev = evaluate_subregion(corr, cum, rho, box, histories=4096)
        _say(f"torso {name}: mean |err| plain {ev['plain_mean']:.4f} -> cnn {ev['cnn_mean']:.4f}; p90 plain "
             f"{ev['plain_p90']:.4f} -> cnn {ev['cnn_p90']:.4f}; within 2% plain {ev['plain_within']:.3f} -> cnn "
             f"{ev['cnn_within']:.3f} ({ev['voxels']} voxels, fallback {ev['fallback']})")
        assert not ev["fallback"] and ev["voxels"] > 200
        res[name] = ev
    assert res["hottest"]["cnn_mean"] < 0.02
    assert res["lung_edge"]["cnn_mean"] < 0.5 * res["lung_edge"]["plain_mean"]
