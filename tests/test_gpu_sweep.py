"""Seeded random-shape GPU parity sweep (through the C ABI), stopping at the first
failing case.

Each case draws a volume (1..200 per axis, at most 2.5 M voxels; a third of the
axes on lengths of the register transforms, 256 among them), a kernel (1..33 per
axis, odd and even, half of them radially symmetric so the real-octant spectrum
runs), the mode (reference circular / centred linear), the algorithm (FFT, or the
direct path when every used extent is <= 16), a density map or none, and a
scale.  The FP64 oracle is kernel_convolution.py:71-74 (reference mode) or the
'same' linear convolution, then scale * rho_w / rho; tolerance max|D - D64| <=
1e-5 max|D64|.  A second sweep puts random volumes through one batch and checks
every member bit-exact against its own single-volume plan, eager and in the graph.  A
third draws volumes with x extents 260..500, whose x lengths take 8-column x tiles.
"""
import sys

import numpy as np
import pytest

from oracle import conv

pytestmark = pytest.mark.gpu

MAX_REL = 1e-5
REGISTER_LENGTHS = [96, 112, 120, 128, 144, 160, 192, 200, 256]


def _say(*a):
    print("[gpu-case]", *a, file=sys.stderr, flush=True)


def _dev(x):
    import torch
    return torch.from_numpy(np.ascontiguousarray(x, dtype=np.float32)).cuda()


def _shape(rng):
    while True:
        s = tuple(int(rng.choice(REGISTER_LENGTHS)) if rng.random() < 0.33 else int(rng.integers(1, 201))
                  for _ in range(3))
        if np.prod(s) <= 2_500_000:
            return s


def _kernel(rng, shape):
    if rng.random() < 0.5:
        k = tuple(int(2 * rng.integers(0, 17) + 1) for _ in range(3))      # odd: radially symmetric
        g = np.meshgrid(*[np.arange(n) - n // 2 for n in k], indexing="ij")
        r = np.sqrt(sum(x.astype(np.float64) ** 2 for x in g))
        return np.exp(-r / rng.uniform(0.5, 4.0)) / (1.0 + r * r)
    k = tuple(int(rng.integers(1, 34)) for _ in range(3))
    return rng.uniform(0.0, 1.0, size=k)


def _cases(seed, n):
    rng = np.random.default_rng(seed)
    for i in range(n):
        shape = _shape(rng)
        kern = _kernel(rng, shape)
        mode = "reference" if rng.random() < 0.6 else "linear"
        ku = [min(k, s) if mode == "reference" else min(k // 2, s - 1) + min(k - 1 - k // 2, s - 1) + 1
              for k, s in zip(kern.shape, shape)]
        algo = "direct" if (max(ku) <= 16 and rng.random() < 0.3) else "fft"
        act = rng.gamma(2.0, 1.0, size=shape) * (rng.random(shape) < 0.3)
        dens = rng.uniform(0.3, 1.9, size=shape) if rng.random() < 0.5 else None
        scale = float(rng.uniform(0.1, 10.0))
        yield i, shape, kern, mode, algo, act, dens, scale


def test_random_shape_sweep_against_the_fp64_oracle():
    from pyvoxeldosimetry_amd import ConvPlan
    worst = 0.0
    for i, shape, kern, mode, algo, act, dens, scale in _cases(2026, 40):
        _say(f"sweep {i}: {shape} kernel {kern.shape} {mode} {algo} density {dens is not None} scale {scale:.3f}")
        plan = ConvPlan(shape, kern.shape, mode=mode, algo=algo)
        plan.set_kernel(kern)
        got = plan.dose(_dev(act), density=_dev(dens) if dens is not None else None,
                        scale=scale).cpu().numpy().astype(np.float64)
        fft_n = plan.fft_n
        plan.close()
        raw = conv.ref_conv_threaded(act, kern, workers=16) if mode == "reference" else conv.linear_same(act, kern)
        exp = conv.epilogue(raw, dens, scale)
        assert np.isfinite(got).all(), (i, shape)
        m = conv.error_metrics(got, exp)
        worst = max(worst, m["max_rel"])
        assert m["max_rel"] <= MAX_REL, (i, shape, kern.shape, mode, algo, fft_n, m)
    _say(f"sweep: worst max_rel {worst:.2e}")


def test_random_long_x_volumes_with_8_column_x_tiles():
    """x extents 260..500 put the x lengths in the fused two-step range 288..512, where the
    plan picks 8-column x tiles: random kernels, modes, densities and scales against the
    FP64 oracle."""
    from pyvoxeldosimetry_amd import ConvPlan
    rng = np.random.default_rng(4242)
    n8 = 0
    for i in range(12):
        shape = (int(rng.integers(260, 501)), int(rng.integers(4, 60)), int(rng.integers(4, 60)))
        kern = _kernel(rng, shape)
        mode = "reference" if rng.random() < 0.6 else "linear"
        act = rng.gamma(2.0, 1.0, size=shape) * (rng.random(shape) < 0.3)
        dens = rng.uniform(0.3, 1.9, size=shape) if rng.random() < 0.5 else None
        scale = float(rng.uniform(0.1, 10.0))
        _say(f"long x {i}: {shape} kernel {kern.shape} {mode}")
        plan = ConvPlan(shape, kern.shape, mode=mode, algo="fft")
        plan.set_kernel(kern)
        got = plan.dose(_dev(act), density=_dev(dens) if dens is not None else None,
                        scale=scale).cpu().numpy().astype(np.float64)
        n8 += plan.tile[0] == 8
        plan.close()
        raw = conv.ref_conv_threaded(act, kern, workers=16) if mode == "reference" else conv.linear_same(act, kern)
        m = conv.error_metrics(got, conv.epilogue(raw, dens, scale))
        assert np.isfinite(got).all() and m["max_rel"] <= MAX_REL, (i, shape, kern.shape, mode, m)
    assert n8 >= 10, n8


def test_random_batch_members_bit_exact_with_single_plans():
    import torch
    from pyvoxeldosimetry_amd import ConvPlan
    from pyvoxeldosimetry_amd.batch import DoseBatch
    rng = np.random.default_rng(77)
    shapes = [_shape(rng) for _ in range(10)]
    kerns = [_kernel(rng, s) for s in shapes]
    times = [1.0, 4.0, 24.0]
    _say("batch sweep shapes", shapes)
    b = DoseBatch(shapes, [torch.from_numpy(k).cuda() for k in kerns], mode="reference")
    acts = []
    for i, s in enumerate(shapes):
        acts.append([_dev(rng.gamma(2.0, 1.0, size=s)) for _ in times])
        b.set_volume_timeseries(i, acts[i], times)
    eager = [o.clone() for o in b.run(graph=False)]
    graph = [o.clone() for o in b.run(graph=True)]
    torch.cuda.synchronize()
    for i, s in enumerate(shapes):
        plan = ConvPlan(s, kerns[i].shape, mode="reference", algo="fft")
        plan.set_kernel(torch.from_numpy(kerns[i]).cuda())
        single = plan.dose_timeseries(acts[i], times)
        torch.cuda.synchronize()
        assert torch.equal(eager[i], single), (i, s, plan.fft_n)
        assert torch.equal(graph[i], single), (i, s, plan.fft_n)
        plan.close()
    b.close()
