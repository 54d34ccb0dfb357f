"""GPU tests of the learned correction: MFMA CNN kernels vs the FP64 oracle, and
invariants of the Monte Carlo.

Tolerances: the GPU multiplies BF16 values exactly and accumulates in FP32, so
layer outputs computed from the same BF16 operands agree with FP64 to ~1e-5
relative (FP32 outputs) or to BF16 rounding (~4e-3, BF16 outputs).  Whole-
network gradients also see BF16 rounding of every stored activation, so they
are compared by relative Frobenius norm (<= 3e-2).
"""
import sys

import numpy as np
import pytest

from oracle import cnn as ocnn

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[gpu-case]", *a, file=sys.stderr, flush=True)


def bf(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(ocnn.bf16_round(a))).to(torch.bfloat16).cuda()


def rel_fro(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


SHAPE = (3, 5, 64)


def _conv(lib, x, cin, w, bias, epi, shape, y=None, act=None, resid=None, y32=None):
    import ctypes
    import torch
    from pyvoxeldosimetry_amd import _native as nat
    p = lambda t: ctypes.c_void_p(t.data_ptr()) if t is not None else ctypes.c_void_p(0)  # noqa: E731
    nat.check(lib.pvd_cnn_conv_fwd(p(x), cin, p(w), p(bias), epi, p(resid), p(y), p(act), p(y32), *shape, None, None,
                                   ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)), "conv")


@pytest.mark.parametrize("cin", [16, 32])
def test_conv_forward_all_epilogues(cin):
    import torch
    from pyvoxeldosimetry_amd import _native as nat
    _say("cnn conv fwd", cin)
    lib = nat.load()
    rng = np.random.default_rng(cin)
    n = int(np.prod(SHAPE))
    x = ocnn.bf16_round(rng.normal(size=SHAPE + (cin,)).astype(np.float32))
    w = ocnn.bf16_round(rng.normal(scale=0.1, size=(27, 32, cin)).astype(np.float32))
    b = rng.normal(scale=0.1, size=32).astype(np.float32)
    z = ocnn.conv3(x, w) + b
    xt, wt, bt = bf(x).reshape(n, cin), bf(w), torch.from_numpy(b).cuda()
    # relu
    y = torch.empty((n, 32), dtype=torch.bfloat16, device="cuda")
    act = torch.empty_like(y)
    _conv(lib, xt, cin, wt, bt, nat.PVD_CNN_EPI_RELU, SHAPE, y=y, act=act)
    got = y.float().cpu().numpy().reshape(SHAPE + (32,))
    np.testing.assert_allclose(got, np.maximum(z, 0), rtol=8e-3, atol=1e-3)
    assert torch.equal(y, act)
    # linear FP32 (channel 0)
    r = torch.empty(n, dtype=torch.float32, device="cuda")
    _conv(lib, xt, cin, wt, bt, nat.PVD_CNN_EPI_LINEAR_F32, SHAPE, y32=r)
    np.testing.assert_allclose(r.cpu().numpy().reshape(SHAPE), z[..., 0], rtol=1e-5, atol=1e-4)
    # accumulate FP32 (no bias)
    acc = torch.ones((n, 32), dtype=torch.float32, device="cuda")
    _conv(lib, xt, cin, wt, None, nat.PVD_CNN_EPI_ACCUM_F32, SHAPE, y32=acc)
    np.testing.assert_allclose(acc.cpu().numpy().reshape(SHAPE + (32,)), 1.0 + ocnn.conv3(x, w), rtol=1e-5, atol=1e-4)
    if cin == 32:
        resid = bf(rng.normal(size=SHAPE + (32,)).astype(np.float32)).reshape(n, 32)
        y2 = torch.empty_like(y)
        _conv(lib, xt, cin, wt, bt, nat.PVD_CNN_EPI_RESID_RELU, SHAPE, y=y2, act=act, resid=resid)
        exp = resid.float().cpu().numpy().reshape(SHAPE + (32,)) + np.maximum(z, 0)
        np.testing.assert_allclose(y2.float().cpu().numpy().reshape(SHAPE + (32,)), exp, rtol=8e-3, atol=2e-3)
        np.testing.assert_allclose(act.float().cpu().numpy().reshape(SHAPE + (32,)), np.maximum(z, 0), rtol=8e-3,
                                   atol=1e-3)
        # inference form: no activations kept, same output bits
        y3 = torch.empty_like(y)
        _conv(lib, xt, cin, wt, bt, nat.PVD_CNN_EPI_RESID_RELU, SHAPE, y=y3, resid=resid)
        assert torch.equal(y2, y3)
        # skip input = the layer's own input: read from the halo in LDS, same bits as
        # the same values passed as a separate tensor (global loads)
        y4, y5 = torch.empty_like(y), torch.empty_like(y)
        _conv(lib, xt, cin, wt, bt, nat.PVD_CNN_EPI_RESID_RELU, SHAPE, y=y4, resid=xt)
        _conv(lib, xt, cin, wt, bt, nat.PVD_CNN_EPI_RESID_RELU, SHAPE, y=y5, resid=xt.clone())
        assert torch.equal(y4, y5)
        np.testing.assert_allclose(y4.float().cpu().numpy().reshape(SHAPE + (32,)), x + np.maximum(z, 0),
                                   rtol=8e-3, atol=2e-3)


@pytest.mark.parametrize("shape", [SHAPE, (35, 37, 64)])
def test_input_layer_with_four_live_channels(shape):
    """cin = 4: a 16-channel input whose channels 4..15 are zero (the packer's layout),
    K packed as 4 taps x 4 channels; the ReLU output matches the oracle and the 16-channel
    kernel on the same input (up to FP32 summation order), and other epilogues or an
    act output are rejected."""
    import torch
    from pyvoxeldosimetry_amd import _native as nat
    _say("cnn input layer cin 4", shape)
    lib = nat.load()
    rng = np.random.default_rng(4)
    n = int(np.prod(shape))
    x = np.zeros(shape + (16,), np.float32)
    x[..., :4] = ocnn.bf16_round(rng.normal(size=shape + (4,)).astype(np.float32))
    w = ocnn.bf16_round(rng.normal(scale=0.1, size=(27, 32, 16)).astype(np.float32))
    b = rng.normal(scale=0.1, size=32).astype(np.float32)
    z = ocnn.conv3(x, w) + b
    xt, wt, bt = bf(x).reshape(n, 16), bf(w), torch.from_numpy(b).cuda()
    y4 = torch.empty((n, 32), dtype=torch.bfloat16, device="cuda")
    y16 = torch.empty_like(y4)
    _conv(lib, xt, 4, wt, bt, nat.PVD_CNN_EPI_RELU, shape, y=y4)
    _conv(lib, xt, 16, wt, bt, nat.PVD_CNN_EPI_RELU, shape, y=y16)
    got = y4.float().cpu().numpy().reshape(shape + (32,))
    np.testing.assert_allclose(got, np.maximum(z, 0), rtol=8e-3, atol=1e-3)
    # BF16 outputs: the two summation orders may round to neighbouring BF16 values
    d = (y4.float() - y16.float()).abs().max().item()
    assert d <= 8e-3 * max(1.0, y16.float().abs().max().item()), d
    act = torch.empty_like(y4)
    with pytest.raises(ValueError):
        _conv(lib, xt, 4, wt, bt, nat.PVD_CNN_EPI_RELU, shape, y=y4, act=act)
    r = torch.empty(n, dtype=torch.float32, device="cuda")
    with pytest.raises(ValueError):
        _conv(lib, xt, 4, wt, bt, nat.PVD_CNN_EPI_LINEAR_F32, shape, y32=r)


@pytest.mark.parametrize("cin", [16, 32])
def test_conv_forward_many_blocks_per_workgroup(cin):
    """A grid with more output blocks than workgroups (ragged in x and y): every
    workgroup walks several blocks, prefetching the next halo during the MFMA loop."""
    import torch
    from pyvoxeldosimetry_amd import _native as nat
    shape = (35, 37, 64)
    _say("cnn conv fwd persistent", cin, shape)
    lib = nat.load()
    rng = np.random.default_rng(7 + cin)
    n = int(np.prod(shape))
    x = ocnn.bf16_round(rng.normal(size=shape + (cin,)).astype(np.float32))
    w = ocnn.bf16_round(rng.normal(scale=0.1, size=(27, 32, cin)).astype(np.float32))
    b = rng.normal(scale=0.1, size=32).astype(np.float32)
    xt, wt, bt = bf(x).reshape(n, cin), bf(w), torch.from_numpy(b).cuda()
    r = torch.empty(n, dtype=torch.float32, device="cuda")
    _conv(lib, xt, cin, wt, bt, nat.PVD_CNN_EPI_LINEAR_F32, shape, y32=r)
    z = ocnn.conv3(x, w) + b
    np.testing.assert_allclose(r.cpu().numpy().reshape(shape), z[..., 0], rtol=1e-5, atol=1e-4)
    y = torch.empty((n, 32), dtype=torch.bfloat16, device="cuda")
    _conv(lib, xt, cin, wt, bt, nat.PVD_CNN_EPI_RELU, shape, y=y)
    np.testing.assert_allclose(y.float().cpu().numpy().reshape(shape + (32,)), np.maximum(z, 0), rtol=8e-3, atol=1e-3)
    if cin == 32:     # residual layer, skip term from the halo (resid == x) and from a copy
        y4, y5 = torch.empty_like(y), torch.empty_like(y)
        _conv(lib, xt, cin, wt, bt, nat.PVD_CNN_EPI_RESID_RELU, shape, y=y4, resid=xt)
        _conv(lib, xt, cin, wt, bt, nat.PVD_CNN_EPI_RESID_RELU, shape, y=y5, resid=xt.clone())
        assert torch.equal(y4, y5)
        np.testing.assert_allclose(y4.float().cpu().numpy().reshape(shape + (32,)), x + np.maximum(z, 0),
                                   rtol=8e-3, atol=2e-3)


@pytest.mark.parametrize("cin", [16, 32])
def test_weight_and_input_gradients(cin):
    import ctypes
    import torch
    from pyvoxeldosimetry_amd import _native as nat
    _say("cnn wgrad/dgrad", cin)
    lib = nat.load()
    rng = np.random.default_rng(100 + cin)
    shape = (5, 3, 64)
    n = int(np.prod(shape))
    x = ocnn.bf16_round(rng.normal(size=shape + (cin,)).astype(np.float32))
    dz = ocnn.bf16_round(rng.normal(size=shape + (32,)).astype(np.float32))
    work = torch.empty(int(lib.pvd_cnn_wgrad_work_floats(*shape)), dtype=torch.float32, device="cuda")
    dw = torch.empty(27 * 32 * cin, dtype=torch.float32, device="cuda")
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    xt, dzt = bf(x), bf(dz)          # keep the device tensors alive across the asynchronous calls
    nat.check(lib.pvd_cnn_wgrad(ctypes.c_void_p(xt.data_ptr()), cin, ctypes.c_void_p(dzt.data_ptr()),
                                ctypes.c_void_p(dw.data_ptr()), ctypes.c_void_p(work.data_ptr()), *shape, st), "wgrad")
    exp = ocnn.conv3_weight_grad(x, dz)
    got = dw.cpu().numpy().reshape(27, 32, cin)
    np.testing.assert_allclose(got, exp, rtol=1e-4, atol=1e-3 * np.abs(exp).max())
    # the same call twice gives identical bits (fixed-order reduction)
    dw2 = torch.empty_like(dw)
    nat.check(lib.pvd_cnn_wgrad(ctypes.c_void_p(xt.data_ptr()), cin, ctypes.c_void_p(dzt.data_ptr()),
                                ctypes.c_void_p(dw2.data_ptr()), ctypes.c_void_p(work.data_ptr()), *shape, st), "wgrad")
    assert torch.equal(dw, dw2)
    if cin == 32:
        w = ocnn.bf16_round(rng.normal(scale=0.1, size=(27, 32, 32)).astype(np.float32))
        wt = torch.empty(27 * 32 * 32, dtype=torch.bfloat16, device="cuda")
        wb = bf(w)
        nat.check(lib.pvd_cnn_flip_weights(ctypes.c_void_p(wb.data_ptr()), ctypes.c_void_p(wt.data_ptr()), st),
                  "flip")
        dx = torch.zeros((n, 32), dtype=torch.float32, device="cuda")
        _conv(lib, dzt.reshape(n, 32), 32, wt, None, nat.PVD_CNN_EPI_ACCUM_F32, shape, y32=dx)
        np.testing.assert_allclose(dx.cpu().numpy().reshape(shape + (32,)), ocnn.conv3_input_grad(dz, w), rtol=1e-4,
                                   atol=1e-4)


def test_network_gradients_match_oracle():
    import torch
    from pyvoxeldosimetry_amd.learned.cnn import ResidualCNN
    _say("cnn full gradients")
    shape = (4, 4, 32)
    n = int(np.prod(shape))
    rng = np.random.default_rng(3)
    net = ResidualCNN(seed=5)
    params = net.params_numpy()
    x = np.zeros(shape + (16,), np.float32)
    x[..., :3] = rng.uniform(0, 1, size=shape + (3,))
    x = ocnn.bf16_round(x)
    dp = rng.uniform(0.5, 1.5, size=shape).astype(np.float32)
    dt = (dp * rng.uniform(0.8, 1.2, size=shape)).astype(np.float32)
    m = (rng.random(shape) > 0.3).astype(np.float32)
    xt = bf(x).reshape(n, 16)
    r = net.forward(xt, shape)
    r_exp, cache = ocnn.forward(params, x.astype(np.float64))
    assert rel_fro(r.cpu().numpy().reshape(shape), r_exp) < 2e-2
    sums = net.loss_and_grad(*(torch.from_numpy(a.reshape(-1)).cuda() for a in (dp, dt, m)), shape)
    loss = (sums[0] / sums[1]).item()
    loss_exp, dr = ocnn.loss_and_grad_r(r_exp, dp, dt, m)
    assert loss == pytest.approx(loss_exp, rel=3e-2)
    net.backward(xt, shape)
    grads = ocnn.backward(params, cache, dr)
    for l in range(6):
        gw = net.grad_w(l).cpu().numpy().reshape(grads[l][0].shape)
        gb = net.grad_b(l).cpu().numpy()
        assert rel_fro(gw, grads[l][0]) < 3e-2, (l, rel_fro(gw, grads[l][0]))
        assert rel_fro(gb, grads[l][1]) < 3e-2, (l, rel_fro(gb, grads[l][1]))


def test_adam_matches_reference_update():
    import torch
    from pyvoxeldosimetry_amd.learned.cnn import ResidualCNN
    _say("adam")
    net = ResidualCNN(seed=1)
    p0 = net.master.cpu().numpy().astype(np.float64)
    g = np.random.default_rng(4).normal(size=p0.shape).astype(np.float32)
    net.grad.copy_(torch.from_numpy(g))
    net.adam_step(lr=1e-2)
    net.grad.copy_(torch.from_numpy(g * 0.5))
    net.adam_step(lr=1e-2)
    m = v = 0.0
    p = p0.copy()
    for t, gg in ((1, g.astype(np.float64)), (2, 0.5 * g.astype(np.float64))):
        m = 0.9 * m + 0.1 * gg
        v = 0.999 * v + 0.001 * gg * gg
        p = p - 1e-2 * (m / (1 - 0.9 ** t)) / (np.sqrt(v / (1 - 0.999 ** t)) + 1e-8)
    np.testing.assert_allclose(net.master.cpu().numpy(), p, rtol=1e-5, atol=1e-6)
    np.testing.assert_array_equal(net.bf16.float().cpu().numpy(), ocnn.bf16_round(net.master.cpu().numpy()))


# ----------------------------------------------------------------- Monte Carlo
def test_mc_energy_conservation_and_escape():
    import torch
    from pyvoxeldosimetry_amd.learned.mc import MonteCarlo
    from pyvoxeldosimetry_amd.learned.spectra import mean_energy
    _say("mc conservation")
    mc = MonteCarlo("Y90")
    rng = np.random.default_rng(0)
    a = torch.from_numpy(rng.uniform(0, 1, size=(20, 18, 16)).astype(np.float32)).cuda()
    rho = torch.from_numpy(rng.choice([0.3, 1.0, 1.85], size=(20, 18, 16)).astype(np.float32)).cuda()
    e = mc.edep(a, rho, 2.0, histories=64, seed=3)
    emitted, deposited, escaped = mc.last_totals.cpu().numpy()
    assert abs(deposited + escaped - emitted) / emitted < 1e-5
    assert abs(e.double().sum().item() - deposited) / deposited < 1e-4
    assert 0.0 < escaped < emitted
    # emitted energy per decay matches the spectrum mean (statistically)
    assert emitted / a.double().sum().item() == pytest.approx(mean_energy("Y90"), rel=1e-2)


def test_mc_density_voxel_scaling_invariance():
    """(rho, voxel) -> (rho/2, 2 voxel) gives the same deposit map (same random numbers)."""
    import torch
    from pyvoxeldosimetry_amd.learned.mc import MonteCarlo
    _say("mc scaling invariance")
    mc = MonteCarlo("Y90")
    rng = np.random.default_rng(1)
    a = torch.from_numpy(rng.uniform(0, 1, size=(16, 16, 16)).astype(np.float32)).cuda()
    rho = torch.from_numpy(rng.uniform(0.5, 2.0, size=(16, 16, 16)).astype(np.float32)).cuda()
    e1 = mc.edep(a, rho, 1.0, histories=16, seed=9).double()
    e2 = mc.edep(a, rho * 0.5, 2.0, histories=16, seed=9).double()
    assert ((e1 - e2).abs().max() / e1.abs().max()).item() < 1e-4


def test_mc_point_kernel_is_isotropic_and_consistent():
    from pyvoxeldosimetry_amd.learned.mc import MonteCarlo
    from pyvoxeldosimetry_amd.learned.spectra import mean_energy
    _say("mc point kernel")
    mc = MonteCarlo("Y90")
    k = mc.point_kernel(2.0, 15, histories=1 << 20)
    assert k.sum() == pytest.approx(mean_energy("Y90"), rel=2e-2)      # nearly nothing escapes 15 voxels
    c = 7
    assert k[c, c, c] == k.max()
    # monotone fall-off along an axis
    prof = k[c, c, c:]
    assert np.all(np.diff(prof[:6]) < 0)


def test_mc_electrons_deposit_nothing_beyond_the_csda_range_of_emax():
    """Electron path lengths are the Katz-Penfold CSDA range R(E0) <= R(E_max) and a
    deposit goes to the voxel where its step starts, so no voxel farther than
    R(E_max) from the emitting voxel receives energy (exact, no statistics)."""
    import torch
    from pyvoxeldosimetry_amd.learned.mc import MonteCarlo
    from pyvoxeldosimetry_amd.learned.spectra import BETA_SPECTRA, katz_range_mm
    _say("mc csda range")
    for nuc, v in (("Y90", 1.0), ("Lu177", 0.25)):
        mc = MonteCarlo(nuc, photons=False)
        n = 31
        a = torch.zeros((n, n, n), dtype=torch.float32, device="cuda")
        c = n // 2
        a[c, c, c] = 1.0
        e = mc.edep(a, None, v, histories=1 << 18, seed=5).cpu().numpy()
        rmax = float(katz_range_mm(BETA_SPECTRA[nuc][0][-1]))
        g = np.abs(np.arange(n) - c)
        gap = np.maximum(g - 1, 0) * v                      # nearest distance between voxel boxes, per axis
        dmin = np.sqrt(gap[:, None, None] ** 2 + gap[None, :, None] ** 2 + gap[None, None, :] ** 2)
        assert e[dmin > rmax].max() == 0.0, nuc
        assert e[dmin > 0.5 * rmax].max() > 0.0, nuc             # the range is actually approached


def test_mc_mean_deposit_per_decay_is_the_spectrum_mean_in_an_infinite_medium():
    """Electrons only, a water cube wider than twice the largest range: nothing escapes
    and the deposited energy per decay is the beta spectrum's mean energy."""
    import torch
    from pyvoxeldosimetry_amd.learned.mc import MonteCarlo
    from pyvoxeldosimetry_amd.learned.spectra import mean_energy
    _say("mc mean deposit")
    for nuc in ("Y90", "Lu177"):
        mc = MonteCarlo(nuc, photons=False)
        a = torch.zeros((15, 15, 15), dtype=torch.float32, device="cuda")
        a[7, 7, 7] = 1.0
        mc.edep(a, None, 2.0, histories=1 << 20, seed=11)
        emitted, deposited, escaped = mc.last_totals.cpu().numpy()
        assert escaped == 0.0, nuc
        assert deposited == pytest.approx(mean_energy(nuc), rel=3e-3), nuc


def _escape_expectation(mu_per_voxel, n, samples=1 << 20, seed=0):
    """E[exp(-mu s)] over isotropic directions from points uniform in the centre voxel
    of an n^3 cube, s the distance to the cube's surface (independent numpy estimate)."""
    rng = np.random.default_rng(seed)
    p = n // 2 + rng.random((samples, 3))
    d = rng.normal(size=(samples, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    with np.errstate(divide="ignore"):
        t = np.where(d > 0, (n - p) / d, np.where(d < 0, -p / d, np.inf))
    return float(np.exp(-mu_per_voxel * t.min(axis=1)).mean())


def test_mc_photon_kerma_escape_and_radial_profile_match_closed_forms():
    """Photons only (one 30 keV line, mu_en = 0.1557 cm^2/g): the escaped fraction equals
    E[exp(-mu_en s_exit)] and the deposit density at radius r follows
    mu_en exp(-mu_en r) / (4 pi r^2) (kerma of a point source)."""
    import torch
    from pyvoxeldosimetry_amd.learned.mc import MonteCarlo
    from pyvoxeldosimetry_amd.learned.spectra import water_mu_en
    _say("mc photon closed forms")
    mc = MonteCarlo("Y90", electrons=False, photon_lines=[(0.03, 1.0)], brems_z=0.0, photon_every=1)
    n, v = 61, 1.0
    a = torch.zeros((n, n, n), dtype=torch.float32, device="cuda")
    c = n // 2
    a[c, c, c] = 1.0
    e = mc.edep(a, None, v, histories=1 << 20, seed=3).cpu().numpy()
    emitted, deposited, escaped = mc.last_totals.cpu().numpy()
    assert emitted == pytest.approx(0.03, rel=1e-6)
    assert deposited + escaped == pytest.approx(emitted, rel=1e-5)
    mu = float(water_mu_en(0.03)) * 0.1 * v                 # per voxel
    assert escaped / emitted == pytest.approx(_escape_expectation(mu, n), abs=2e-3)
    g = np.arange(n) - c
    r = np.sqrt(g[:, None, None] ** 2 + g[None, :, None] ** 2 + g[None, None, :] ** 2) * v
    for r0 in (8.0, 14.0, 20.0):
        shell = (r > r0 - 1.0) & (r <= r0 + 1.0)
        got = e[shell].sum() / (shell.sum() * v ** 3)
        exp = 0.03 * (mu / v) * (np.exp(-(mu / v) * r[shell]) / (4 * np.pi * r[shell] ** 2)).mean()
        assert got == pytest.approx(exp, rel=0.05), r0


def test_mc_lu177_gamma_lines_and_y90_bremsstrahlung_energy_budget():
    """With photons, emitted energy per decay = beta mean + sum of line energy x yield
    (Lu-177), or stays the beta mean with Y(E0) of it carried by bremsstrahlung (Y-90);
    deposited + escaped = emitted for every configuration."""
    import torch
    from pyvoxeldosimetry_amd.learned.mc import MonteCarlo
    from pyvoxeldosimetry_amd.learned.spectra import PHOTON_LINES, mean_energy
    _say("mc photon budget")
    rng = np.random.default_rng(2)
    a = torch.from_numpy(rng.uniform(0, 1, size=(24, 20, 18)).astype(np.float32)).cuda()
    rho = torch.from_numpy(rng.choice([0.3, 1.0, 1.85], size=(24, 20, 18)).astype(np.float32)).cuda()
    decays = a.double().sum().item()
    for nuc in ("Lu177", "Y90"):
        mc = MonteCarlo(nuc)
        mc.edep(a, rho, 2.0, histories=256, seed=4)
        emitted, deposited, escaped = mc.last_totals.cpu().numpy()
        assert abs(deposited + escaped - emitted) / emitted < 1e-5, nuc
        want = mean_energy(nuc) + sum(e * y for e, y in PHOTON_LINES[nuc])
        assert emitted / decays == pytest.approx(want, rel=1.5e-2), nuc
        nophot = MonteCarlo(nuc, photons=False)
        e_np = nophot.edep(a, rho, 2.0, histories=256, seed=4).sum().item()
        if nuc == "Lu177":
            assert e_np < deposited                            # the gamma lines add deposits
        else:                                                  # bremsstrahlung carries ~0.4 % away
            assert 0.99 * e_np < deposited < e_np


# ----------------------------------------------------------------- learned pipeline
def test_correction_fallback_rule():
    """Out-of-bound corrections injected from the host fall back to the plain dose."""
    import torch
    from pyvoxeldosimetry_amd.learned.pipeline import LearnedCorrector, synth_phantom
    _say("fallback")
    corr = LearnedCorrector("Y90", voxel_mm=2.0, kernel_size=11, kernel_histories=1 << 20)
    act, rho = synth_phantom(5, (32, 32, 32), corr.device)
    dp = corr.plain(act, rho).clone()
    r = torch.full(act.shape, 0.1, device="cuda")
    c = corr.correct(act, rho, r_override=r, region=False)
    assert not c.fallback and c.exceed == 0
    torch.testing.assert_close(c.dose, dp * 1.1, rtol=1e-6, atol=0)
    r_bad = r.clone()
    r_bad[0, 0, 0] = 7.0                                   # one voxel out of bound: voxelwise fallback
    c = corr.correct(act, rho, r_override=r_bad, region=False)
    assert c.dose[0, 0, 0].item() == pytest.approx(dp[0, 0, 0].item())
    r_bad = torch.full(act.shape, -3.0, device="cuda")     # everything out of bound: whole-volume fallback
    c = corr.correct(act, rho, r_override=r_bad, region=False)
    assert c.fallback and c.exceed > 0
    torch.testing.assert_close(c.dose, dp)


def test_input_maxima_equal_torch_max():
    """pvd_cnn_input_maxima gives max(activity) and max(dose) exactly (a maximum has no
    rounding): tiny, ragged, all-negative and more-than-one-voxel-per-thread sizes."""
    import ctypes
    import torch
    from pyvoxeldosimetry_amd import _native as nat
    _say("input maxima")
    lib = nat.load()
    g = torch.Generator(device="cuda").manual_seed(11)
    for n in [1, 255, 70001, 2048 * 1024 + 7, 5_000_000]:
        a = torch.randn(n, device="cuda", generator=g) * 3.0
        d = -torch.rand(n, device="cuda", generator=g) - 0.5          # all negative
        if n > 1:
            a[n // 3] = 1e6                                           # one spike inside
        m = torch.full((nat.PVD_CNN_MAXIMA_FLOATS,), float("nan"), device="cuda")
        nat.check(lib.pvd_cnn_input_maxima(ctypes.c_void_p(a.data_ptr()), ctypes.c_void_p(d.data_ptr()), n,
                                           ctypes.c_void_p(m.data_ptr()), None))
        assert m[0].item() == a.max().item(), n
        assert m[1].item() == d.max().item(), n
    with pytest.raises(ValueError):
        nat.check(lib.pvd_cnn_input_maxima(None, None, 0, None, None))


def test_device_fallback_decision_at_its_boundary():
    """pvd_cnn_apply_fallback copies the plain dose exactly when exceed > frac * region
    (the host rule of Correction.fallback), with no host synchronisation."""
    import ctypes
    import torch
    from pyvoxeldosimetry_amd import _native as nat
    from pyvoxeldosimetry_amd.learned.pipeline import Correction
    _say("device fallback")
    lib = nat.load()
    n = 70001                                               # not a multiple of the block size
    dp = torch.rand(n, device="cuda") + 1.0
    for ex, reg, frac in [(5, 100, 0.05), (6, 100, 0.05), (0, 0, 0.05), (1, 0, 0.05), (0, 7, 0.0), (1, 7, 0.0)]:
        counts = torch.tensor([ex, reg], dtype=torch.int64, device="cuda")
        out = torch.full((n,), -1.0, device="cuda")
        nat.check(lib.pvd_cnn_apply_fallback(ctypes.c_void_p(dp.data_ptr()), ctypes.c_void_p(counts.data_ptr()),
                                             ctypes.c_double(frac), ctypes.c_void_p(out.data_ptr()), n, None))
        expect = Correction(out, dp, None, counts, frac).fallback
        assert expect == (reg > 0 and ex > frac * reg)
        if expect:
            assert torch.equal(out, dp), (ex, reg, frac)
        else:
            assert bool((out == -1.0).all()), (ex, reg, frac)


def test_short_training_fits_its_phantoms():
    """A short run of the GPU training loop (forward, loss, backward, Adam, all on
    the HIP kernels) lowers the loss and the error against its own Monte Carlo targets."""
    import torch
    from pyvoxeldosimetry_amd.learned.pipeline import LearnedCorrector, voxel_errors
    from pyvoxeldosimetry_amd.learned.train import build_training_set, train
    _say("short training 64^3")
    torch.manual_seed(0)
    shape = (64, 64, 64)
    corr = LearnedCorrector("Y90", voxel_mm=2.0, kernel_size=13, kernel_histories=1 << 22, bound=0.9, seed=1)
    data = build_training_set(corr, 24, shape, histories=1024)

    def mean_err(use_cnn):
        e = []
        for act, rho, tgt in data[:6]:
            c = corr.correct(act, rho, region=False)
            e.append(voxel_errors(c.dose if use_cnn else c.plain, tgt).mean().item())
        return sum(e) / len(e)

    plain = mean_err(False)
    assert mean_err(True) == pytest.approx(plain, rel=1e-6)     # untrained: r = 0
    hist, _ = train(corr, data, epochs=12, lr=1e-3)
    fitted = mean_err(True)
    _say(f"loss per epoch {['%.3e' % h for h in hist]}; training-phantom error plain {plain:.4f} -> cnn {fitted:.4f}")
    assert hist[-1] < 0.5 * hist[0]
    assert fitted < plain


def test_gpu_training_from_scratch_beats_plain_on_held_out_monte_carlo():
    """From scratch on the GPU (forward, loss, backward and Adam on the HIP kernels):
    128 seeded 32^3 phantoms with Monte Carlo targets (1024 histories), 60 epochs; on
    8 held-out phantoms (seeds 20000..20007) against 4096-history Monte Carlo, the
    corrected dose's mean voxel error (>= 10 %-max region) is below the plain dose's."""
    import torch
    from pyvoxeldosimetry_amd.learned.pipeline import LearnedCorrector, synth_phantom, voxel_errors
    from pyvoxeldosimetry_amd.learned.train import build_training_set, train
    _say("training from scratch, held-out evaluation")
    torch.manual_seed(0)
    shape = (32, 32, 32)
    corr = LearnedCorrector("Y90", voxel_mm=2.0, kernel_size=13, kernel_histories=1 << 22, bound=0.9, seed=1)
    data = build_training_set(corr, 128, shape, histories=1024)
    hist, _ = train(corr, data, epochs=60, lr=2e-3)
    del data
    plain, cnn = [], []
    for s in range(20000, 20008):
        act, rho = synth_phantom(s, shape, corr.device)
        ref = corr.mc_dose(act, rho, 4096, seed=s)
        c = corr.correct(act, rho)
        plain.append(voxel_errors(c.plain, ref).mean().item())
        cnn.append(voxel_errors(c.dose, ref).mean().item())
    p, q = sum(plain) / len(plain), sum(cnn) / len(cnn)
    _say(f"loss {hist[0]:.3e} -> {hist[-1]:.3e}; held-out mean |err| plain {p:.4f} -> cnn {q:.4f}")
    assert hist[-1] < hist[0]
    assert q < 0.9 * p


def test_pretrained_correction_beats_plain_on_held_out_monte_carlo():
    """The shipped weights (learned/weights: trained at 64^3 on seeds 10000..11199, then
    at 128^3 on seeds 10000..11999) against Monte Carlo of held-out phantoms (seeds
    20000..): the corrected dose is much
    closer to Monte Carlo than the plain density-scaled convolution."""
    from pyvoxeldosimetry_amd.learned.pipeline import EVAL_SEEDS_TEST, LearnedCorrector, evaluate
    _say("pretrained vs held-out MC")
    corr = LearnedCorrector.pretrained("Y90", 2.0)
    ev = evaluate(corr, EVAL_SEEDS_TEST, (64, 64, 64), histories=8192)
    _say(f"held-out: mean |err| plain {ev['plain_mean']:.4f} -> cnn {ev['cnn_mean']:.4f}; p90 plain "
         f"{ev['plain_p90']:.4f} -> cnn {ev['cnn_p90']:.4f}; within 2% plain {ev['plain_within']:.3f} -> cnn "
         f"{ev['cnn_within']:.3f}; fallbacks {ev['fallbacks']}")
    assert ev["fallbacks"] == 0
    assert ev["cnn_mean"] < 0.5 * ev["plain_mean"]
    assert ev["cnn_p90"] < 0.5 * ev["plain_p90"]


def test_plain_dose_matches_monte_carlo_in_water():
    """Homogeneous water: the convolution with the Monte Carlo point kernel reproduces
    the Monte Carlo (validates the kernel and the density epilogue at rho = 1)."""
    import torch
    from pyvoxeldosimetry_amd.learned.pipeline import LearnedCorrector, synth_phantom, voxel_errors
    _say("plain vs MC in water")
    corr = LearnedCorrector("Y90", voxel_mm=2.0, kernel_size=13, kernel_histories=1 << 23)
    act, _ = synth_phantom(7, (48, 48, 64), corr.device)
    rho = torch.ones_like(act)
    ref = corr.mc_dose(act, rho, histories=2048, seed=3)
    ref2 = corr.mc_dose(act, rho, histories=2048, seed=4)
    plain = corr.plain(act, rho)
    region = ref >= 0.1 * ref.max()
    rel = voxel_errors(plain, ref)
    floor = voxel_errors(ref2, ref)                # two independent MC runs: sqrt(2) x the noise
    bias = ((plain[region] - ref[region]) / ref[region]).mean().item()
    _say(f"water: plain-vs-MC mean |err| {rel.mean().item():.4f}, MC-vs-MC {floor.mean().item():.4f}, "
         f"signed bias {bias:+.4f}")
    assert abs(bias) < 0.005
    assert rel.mean().item() < 0.85 * floor.mean().item()     # noise-limited: ~1/sqrt(2) of the MC-MC spread
