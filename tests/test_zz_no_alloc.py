"""The allocation-free direct dose entry point (pvd_conv_dose_direct_noalloc).

Checked through the C ABI against the plan-based direct path it mirrors
(pvd_plan_create(algo=direct) + pvd_plan_set_kernel + pvd_conv_dose): two calls
in a row, the second on a smaller grid, in both convolution modes, bit-exact,
and against the FP64 oracle within 1e-5 of the maximum dose.
"""
import ctypes
import sys

import numpy as np
import pytest

from oracle import conv, dpk


def _call(lib, shape, kshape, mode, k_dev, a_dev, d_dev, out_dev, scale, stream):
    n = (ctypes.c_int * 3)(*shape)
    k = (ctypes.c_int * 3)(*kshape)
    return lib.pvd_conv_dose_direct_noalloc(n, k, mode, ctypes.c_void_p(k_dev), ctypes.c_void_p(a_dev),
                                            ctypes.c_void_p(d_dev), ctypes.c_void_p(out_dev), scale,
                                            ctypes.c_void_p(stream))


def test_oversize_kernel_is_rejected_before_any_device_work():
    from pyvoxeldosimetry_amd import _native as nat
    lib = nat.load()
    fake = 256      # never dereferenced: the extent check comes first
    rc = _call(lib, (64, 64, 64), (17, 17, 17), nat.PVD_MODE_REFERENCE, fake, fake, 0, fake, 1.0, 0)
    assert rc == nat.PVD_ENOTSUP
    assert b"exceeds 16" in lib.pvd_last_error()
    # an unknown mode is an argument error
    rc = _call(lib, (64, 64, 64), (3, 3, 3), 7, fake, fake, 0, fake, 1.0, 0)
    assert rc == nat.PVD_EINVAL


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["reference", "linear"])
def test_noalloc_matches_the_plan_direct_path_on_two_calls(mode):
    import torch
    from pyvoxeldosimetry_amd import ConvPlan
    from pyvoxeldosimetry_amd import _native as nat
    lib = nat.load()
    rng = np.random.default_rng(71)
    stream = torch.cuda.current_stream().cuda_stream
    cases = [((48, 40, 36), dpk.y90_kernel(4.42, (11, 11, 11), fix_centre=True), True),
             ((20, 17, 13), dpk.lu177_kernel(2.0, (9, 7, 5)), False)]
    for shape, k, with_density in cases:          # the second call is smaller
        print(f"[gpu-case] noalloc {mode} {shape} kernel {k.shape}", file=sys.stderr, flush=True)
        act = rng.random(shape).astype(np.float32)
        dens = rng.uniform(0.3, 1.9, size=shape).astype(np.float32) if with_density else None
        a_d = torch.from_numpy(act).cuda()
        d_d = torch.from_numpy(dens).cuda() if dens is not None else None
        k_d = torch.from_numpy(k).cuda()
        out = torch.full(shape, float("nan"), dtype=torch.float32, device="cuda")
        rc = _call(lib, shape, k.shape, nat.MODES[mode], k_d.data_ptr(), a_d.data_ptr(),
                   d_d.data_ptr() if d_d is not None else 0, out.data_ptr(), 0.75, stream)
        nat.check(rc, "pvd_conv_dose_direct_noalloc")
        plan = ConvPlan(shape, k.shape, mode=mode, algo="direct")
        plan.set_kernel(k)
        ref = plan.dose(a_d, density=d_d, scale=0.75)
        torch.cuda.synchronize()
        plan.close()
        np.testing.assert_array_equal(out.cpu().numpy(), ref.cpu().numpy())
        fn = conv.ref_conv if mode == "reference" else conv.linear_same
        exp = conv.epilogue(fn(act.astype(np.float64), k), dens) * 0.75
        assert conv.error_metrics(out.cpu().numpy(), exp)["max_rel"] <= 1e-5
