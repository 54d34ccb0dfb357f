"""GPU test at a size where 32-bit byte offsets no longer cover a 256-point X pass.

256 x 2048 x 2048 (reference mode, 3^3 kernel): the workspace x-planes are
2048 x 1040 complex apart, so the last planes lie more than 4 GiB past the first.
The plan must route X through the 64-bit generic pass (the register X passes
address with 32-bit byte offsets); the check is the
size-independent closed form of the reference convolution for a few deltas,
D[(p + u) mod N] = a K[u] (kernel_convolution.py:71-74 with a kernel anchored at
its origin), and zero elsewhere.
"""
import sys

import numpy as np
import pytest

from oracle import dpk

pytestmark = pytest.mark.gpu


def test_deltas_at_256x2048x2048_land_where_the_closed_form_puts_them():
    import torch
    from pyvoxeldosimetry_amd import ConvPlan
    shape = (256, 2048, 2048)
    print("[gpu-case] 256x2048x2048 deltas", file=sys.stderr, flush=True)
    k = dpk.lu177_kernel(2.0, (3, 3, 3))
    plan = ConvPlan(shape, k.shape, mode="reference", algo="fft")
    assert plan.fft_n == shape and plan.fft_path[0] == "stockham"
    plan.set_kernel(k)
    act = torch.zeros(shape, dtype=torch.float32, device="cuda")
    deltas = [((253, 2040, 2000), 1.0), ((3, 100, 7), 0.5), ((130, 1500, 1024), 2.0), ((255, 2047, 2047), 1.5)]
    for p, a in deltas:
        act[p] = a
    out = plan.dose(act)
    del act
    plan.close()
    kt = torch.from_numpy(k.astype(np.float32)).cuda()
    peak = max(a for _, a in deltas) * float(k.max())
    for p, a in deltas:
        idx = [torch.arange(p[d], p[d] + 3, device="cuda") % shape[d] for d in range(3)]
        block = out[idx[0][:, None, None], idx[1][None, :, None], idx[2][None, None, :]]
        assert float((block - a * kt).abs().max()) <= 1e-5 * peak, p
        out[idx[0][:, None, None], idx[1][None, :, None], idx[2][None, None, :]] = 0.0
    assert float(out.abs().max()) <= 1e-5 * peak
