"""GPU tests of the batched path (one launch per pass over uneven volumes, hipGraph replay)."""
import sys

import numpy as np
import pytest

from oracle import conv, dpk

pytestmark = pytest.mark.gpu

MAX_REL = 1e-5


def _say(*a):
    print("[gpu-case]", *a, file=sys.stderr, flush=True)


def dev(x):
    import torch
    return torch.from_numpy(np.ascontiguousarray(x, dtype=np.float32)).cuda()


SHAPES = [(40, 36, 32), (23, 29, 31), (64, 48, 50), (17, 61, 20), (50, 50, 44)]


def make_batch(mode="reference"):
    rng = np.random.default_rng(33)
    kernels = []
    for i, _ in enumerate(SHAPES):
        if i % 2 == 0:
            kernels.append(dpk.y90_kernel(4.0, (11, 11, 11), fix_centre=True))
        else:
            kernels.append(dpk.lu177_kernel(1.0, (9, 9, 9)))
    acts = [[rng.random(s).astype(np.float32) for _ in range(4)] for s in SHAPES]
    dens = [rng.uniform(0.3, 1.9, size=s).astype(np.float32) if i % 3 == 0 else None for i, s in enumerate(SHAPES)]
    return kernels, acts, dens


def test_batch_matches_single_plans_bit_exact_and_oracle():
    from pyvoxeldosimetry_amd import ConvPlan
    from pyvoxeldosimetry_amd.batch import DoseBatch
    _say("batch vs single plans")
    kernels, acts, dens = make_batch()
    t = [4.0, 24.0, 72.0, 144.0]
    b = DoseBatch(SHAPES, kernels, mode="reference")
    for i in range(len(SHAPES)):
        b.set_volume_timeseries(i, [dev(a) for a in acts[i]], t, density=dev(dens[i]) if dens[i] is not None else None)
    for graph in (False, True):
        outs = [o.cpu().numpy() for o in b.run(graph=graph)]
        for i, shp in enumerate(SHAPES):
            plan = ConvPlan(shp, kernels[i].shape, mode="reference", algo="fft")
            plan.set_kernel(kernels[i])
            single = plan.dose_timeseries([dev(a) for a in acts[i]], t,
                                          density=dev(dens[i]) if dens[i] is not None else None).cpu().numpy()
            np.testing.assert_array_equal(outs[i], single, err_msg=f"volume {i} graph={graph}")
            exp = conv.epilogue(conv.ref_absorbed_dose([a.astype(np.float64) for a in acts[i]], t, kernels[i]),
                                dens[i])
            assert conv.error_metrics(outs[i], exp)["max_rel"] <= MAX_REL, i
    b.close()


def test_graph_replays_track_new_inputs():
    import torch
    from pyvoxeldosimetry_amd.batch import DoseBatch
    _say("graph replays")
    kernels, acts, dens = make_batch("linear")
    b = DoseBatch(SHAPES, kernels, mode="linear")
    rng = np.random.default_rng(5)
    for rep in range(3):
        inputs = [dev(rng.random(s)) for s in SHAPES]
        for i in range(len(SHAPES)):
            b.set_volume(i, [inputs[i]], [1.0 + rep])
        graph_out = [o.clone() for o in b.run(graph=True)]
        torch.cuda.synchronize()
        _ = torch.randn(1000, device="cuda").sum().item()       # eager work between replays
        eager_out = [o.clone() for o in b.run(graph=False)]
        torch.cuda.synchronize()
        for i in range(len(SHAPES)):
            assert torch.equal(graph_out[i], eager_out[i]), (rep, i)
            exp = conv.linear_same(inputs[i].cpu().numpy().astype(np.float64), kernels[i]) * (1.0 + rep)
            assert conv.error_metrics(graph_out[i].cpu().numpy(), exp)["max_rel"] <= MAX_REL
    b.close()


def test_batch_with_256_point_volume_matches_single_plans():
    """A 256x256 volume next to smaller ones: the batch launch uses the register-resident
    256-point path for that volume and the generic path for the others, bit-exact with
    single-volume plans."""
    from pyvoxeldosimetry_amd import ConvPlan
    from pyvoxeldosimetry_amd.batch import DoseBatch
    _say("batch with a 256-point volume")
    shapes = [(256, 256, 24), (40, 36, 32), (64, 64, 30)]
    rng = np.random.default_rng(44)
    kernels = [dpk.lu177_kernel(2.0, (9, 9, 9)), dpk.y90_kernel(4.0, (11, 11, 11), fix_centre=True),
               dpk.lu177_kernel(2.0, (9, 9, 9))]
    acts = [rng.random(s).astype(np.float32) for s in shapes]
    b = DoseBatch(shapes, kernels, mode="reference")
    for i in range(len(shapes)):
        b.set_volume(i, [dev(acts[i])], [1.0])
    outs = [o.cpu().numpy() for o in b.run(graph=True)]
    for i, shp in enumerate(shapes):
        plan = ConvPlan(shp, kernels[i].shape, mode="reference", algo="fft")
        plan.set_kernel(kernels[i])
        np.testing.assert_array_equal(outs[i], plan.dose(dev(acts[i])).cpu().numpy(), err_msg=f"volume {i}")
        assert conv.error_metrics(outs[i], conv.ref_conv(acts[i].astype(np.float64), kernels[i]))["max_rel"] <= MAX_REL
    b.close()


def test_batch_with_two_step_length_volume():
    """A volume whose y length is a register two-step length (560 = 16 x 35) next to
    others: in the batch launch it takes the generic stages, alone the two-step; both
    match the FP64 oracle and each other within FP32 rounding."""
    from pyvoxeldosimetry_amd import ConvPlan
    from pyvoxeldosimetry_amd.batch import DoseBatch
    _say("batch with a 16*B-length volume")
    shapes = [(6, 560, 12), (40, 36, 32)]
    rng = np.random.default_rng(45)
    kernels = [dpk.lu177_kernel(2.0, (3, 9, 3)), dpk.y90_kernel(4.0, (11, 11, 11), fix_centre=True)]
    acts = [rng.random(s).astype(np.float32) for s in shapes]
    b = DoseBatch(shapes, kernels, mode="reference")
    for i in range(len(shapes)):
        b.set_volume(i, [dev(acts[i])], [1.0])
    outs = [o.cpu().numpy() for o in b.run(graph=True)]
    for i, shp in enumerate(shapes):
        exp = conv.ref_conv(acts[i].astype(np.float64), kernels[i])
        plan = ConvPlan(shp, kernels[i].shape, mode="reference", algo="fft")
        plan.set_kernel(kernels[i])
        single = plan.dose(dev(acts[i])).cpu().numpy()
        plan.close()
        assert conv.error_metrics(outs[i], exp)["max_rel"] <= MAX_REL, f"batch volume {i}"
        assert conv.error_metrics(single, exp)["max_rel"] <= MAX_REL, f"single volume {i}"
        assert np.abs(outs[i] - single).max() <= 1e-6 * np.abs(exp).max()
    b.close()


def test_batch_errors():
    from pyvoxeldosimetry_amd.batch import DoseBatch
    _say("batch errors")
    kernels, acts, _ = make_batch()
    b = DoseBatch(SHAPES[:2], kernels[:2])
    with pytest.raises(ValueError, match="no inputs"):
        b.run()
    with pytest.raises(ValueError):
        b.set_volume(5, [dev(acts[0][0])], [1.0])
    with pytest.raises(ValueError):
        b.set_volume(0, [dev(acts[1][0])], [1.0])          # wrong shape for volume 0
    b.close()
