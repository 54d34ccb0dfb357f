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
    """A 256x256 volume next to smaller ones: the 256-point volume runs the register-resident
    path and the others theirs, bit-exact with single-volume plans."""
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


def _batch_vs_single(shapes, kernels, acts, tile=None, mode="reference", smooth_lengths=False):
    """Run the volumes as one batch (graph) and as single plans with the same tile; check
    every batch output bit-exact against its single plan and within 1e-5 of the oracle."""
    from pyvoxeldosimetry_amd import ConvPlan
    from pyvoxeldosimetry_amd.batch import DoseBatch
    b = DoseBatch(shapes, kernels, mode=mode, tile=tile, smooth_lengths=smooth_lengths)
    for i in range(len(shapes)):
        b.set_volume(i, [dev(acts[i])], [1.0])
    outs = [o.cpu().numpy() for o in b.run(graph=True)]
    paths = []
    for i, shp in enumerate(shapes):
        ref = conv.ref_conv if mode == "reference" else conv.linear_same
        exp = ref(acts[i].astype(np.float64), kernels[i])
        plan = ConvPlan(shp, kernels[i].shape, mode=mode, algo="fft", tile=tile, smooth_lengths=smooth_lengths)
        plan.set_kernel(kernels[i])
        single = plan.dose(dev(acts[i])).cpu().numpy()
        paths.append(plan.fft_path)
        plan.close()
        assert conv.error_metrics(outs[i], exp)["max_rel"] <= MAX_REL, f"batch volume {i} {shp}"
        np.testing.assert_array_equal(outs[i], single, err_msg=f"volume {i} {shp}")
    b.close()
    return paths


def test_batch_with_two_step_length_volume():
    """A volume whose y length is a register two-step length (560 = 16 x 35) next to
    others: the batch runs it through the same two-step instance as its own plan."""
    _say("batch with a 16*B-length volume")
    shapes = [(6, 560, 12), (40, 36, 32)]
    rng = np.random.default_rng(45)
    kernels = [dpk.lu177_kernel(2.0, (3, 9, 3)), dpk.y90_kernel(4.0, (11, 11, 11), fix_centre=True)]
    acts = [rng.random(s).astype(np.float32) for s in shapes]
    paths = _batch_vs_single(shapes, kernels, acts)
    assert paths[0][1] == "two_step"


@pytest.mark.parametrize("tile", [(16, 16, 16), (32, 32, 16), (8, 8, 8)])
def test_batch_two_step_and_generic_lengths_at_every_tile(tile):
    """y and x lengths of 560 and 320 (two-step) and 375 (generic Stockham with the
    7-smooth lengths) at 8-, 16- and 32-line tiles: the Stockham stages need
    lines * P <= 16 * block, which a batch launch must keep for every member
    (16 x 560 = 8960 and 32 x 375 = 12000 items)."""
    _say("batch tiles", tile)
    shapes = [(6, 560, 12), (560, 6, 12), (5, 320, 9), (320, 7, 9), (9, 375, 7), (375, 5, 8)]
    rng = np.random.default_rng(46)
    kernels = [dpk.lu177_kernel(2.0, (3, 5, 3)) for _ in shapes]
    acts = [rng.random(s).astype(np.float32) for s in shapes]
    paths = _batch_vs_single(shapes, kernels, acts, tile=tile, smooth_lengths=True)
    assert paths[0][1] == paths[2][1] == "two_step" and paths[1][0] == paths[3][0] == "two_step"
    assert paths[4][1] == paths[5][0] == "stockham"


def test_batch_256_point_volume_keeps_its_register_path_beside_a_640_length_neighbour():
    """The 256-point path runs only with 256-thread blocks; a neighbour whose y pass
    needs a larger block (640 rows) runs in a launch of its own."""
    _say("batch 256 beside 640")
    shapes = [(256, 256, 24), (8, 640, 10)]
    rng = np.random.default_rng(47)
    kernels = [dpk.lu177_kernel(2.0, (9, 9, 9)), dpk.lu177_kernel(2.0, (3, 3, 3))]
    acts = [rng.random(s).astype(np.float32) for s in shapes]
    paths = _batch_vs_single(shapes, kernels, acts)
    assert paths[0][:2] == ("sq256", "sq256")


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
