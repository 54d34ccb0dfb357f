"""Learned heterogeneity correction of the kernel-convolution dose.

plain dose     D_p = conv(A, K_mc) * rho_w / rho      (HIP FFT path, linear mode)
               K_mc: Monte Carlo point kernel of the nuclide in water (same
               transport model as the targets, so water is consistent)
correction     r = CNN(A / max A, rho, D_p / max D_p)  (BF16 MFMA kernels)
corrected      D = D_p (1 + r) where |r| <= bound, D_p elsewhere (optionally only where
               D_p >= region_frac * max D_p, the network then running only on the
               blocks within its receptive field of that region).  If more than
               `max_exceed_frac` of the >= 10 %-max voxels exceed the bound, the whole
               volume falls back to D_p (reported in the result)
targets        Monte Carlo dose of the same activity in the heterogeneous medium
"""
from __future__ import annotations

import ctypes
import math
import os
import time

import numpy as np

from .. import _native as nat
from ..plan import ConvPlan
from .cnn import ResidualCNN
from .mc import MonteCarlo

WEIGHTS_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "weights")
PRETRAINED = {("Y90", 2.0): "cnn_y90_2mm.safetensors", ("Lu177", 0.98): "cnn_lu177_0p98mm.safetensors"}
EVAL_SEEDS_TEST = range(20000, 20006)      # held-out phantoms (training used seeds 10000..11999)


def synth_phantom(seed: int, shape=(128, 128, 128), device=None):
    """Seeded heterogeneous patch: soft tissue with lung and bone inserts, a
    background uptake and spherical lesions (activity in arbitrary units)."""
    import torch
    dev = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
    rng = np.random.default_rng(seed)
    n = np.array(shape, dtype=np.float64)
    x = torch.arange(shape[0], device=dev, dtype=torch.float32).view(-1, 1, 1)
    y = torch.arange(shape[1], device=dev, dtype=torch.float32).view(1, -1, 1)
    z = torch.arange(shape[2], device=dev, dtype=torch.float32).view(1, 1, -1)

    def ellipsoid(c, s):
        return ((x - c[0]) / s[0]) ** 2 + ((y - c[1]) / s[1]) ** 2 + ((z - c[2]) / s[2]) ** 2 <= 1.0

    rho = torch.full(shape, float(rng.uniform(0.98, 1.06)), device=dev)
    lung = torch.zeros(shape, dtype=torch.bool, device=dev)
    bone = torch.zeros(shape, dtype=torch.bool, device=dev)
    for _ in range(int(rng.integers(1, 3))):
        m = ellipsoid(rng.uniform(0.1, 0.9, 3) * n, rng.uniform(0.12, 0.35, 3) * n)
        rho[m] = float(rng.uniform(0.2, 0.4))
        lung |= m
    for _ in range(int(rng.integers(0, 3))):
        m = ellipsoid(rng.uniform(0.1, 0.9, 3) * n, rng.uniform(0.04, 0.15, 3) * n)
        rho[m] = float(rng.uniform(1.5, 1.9))
        bone |= m
    act = torch.full(shape, float(rng.uniform(0.03, 0.12)), device=dev)
    act[lung] *= 0.3
    act[bone] *= 0.1
    for _ in range(int(rng.integers(2, 6))):
        c = rng.uniform(0.1, 0.9, 3) * n
        r = float(rng.uniform(2.0, 0.1 * n.min()))
        act[ellipsoid(c, (r, r, r))] = float(rng.uniform(0.5, 1.0))
    return act.contiguous(), rho.contiguous()


class Correction:
    """Result of `LearnedCorrector.correct`: ``dose`` (corrected, FP32), ``plain`` (the
    plain-convolution dose), ``r`` (predicted relative correction).  ``exceed`` (voxels
    >= 10 % of max whose |r| passed the bound) and ``fallback`` (the whole volume fell
    back to the plain dose) read the device counts on first access, one host
    synchronisation; the fallback itself was applied on the device."""

    def __init__(self, dose, plain, r, counts, max_exceed_frac: float):
        self.dose, self.plain, self.r = dose, plain, r
        self._counts, self._frac, self._host = counts, max_exceed_frac, None

    def _read(self):
        if self._host is None:
            self._host = tuple(int(v) for v in self._counts.tolist())
        return self._host

    @property
    def exceed(self) -> int:
        return self._read()[0]

    @property
    def fallback(self) -> bool:
        n_ex, n_region = self._read()
        return n_region > 0 and n_ex > self._frac * n_region


class LearnedCorrector:
    def __init__(self, nuclide: str = "Y90", voxel_mm: float = 2.0, kernel_size: int = 13,
                 kernel_histories: int = 1 << 23, bound: float = 0.5, max_exceed_frac: float = 0.05, seed: int = 0,
                 device=None, region_frac: float = 0.05, photons: bool = True):
        """``photons``: the Monte Carlo (point kernel and targets) includes the nuclide's
        photons (gamma lines, bremsstrahlung); weights record the setting they were
        trained with."""
        import torch
        nat.require_device()
        self.lib = nat.load()
        self.device = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
        self.voxel_mm = float(voxel_mm)
        self.nuclide = nuclide
        self.photons = bool(photons)
        self.mc = MonteCarlo(nuclide, device=self.device, photons=self.photons)
        k = self.mc.point_kernel(voxel_mm, kernel_size, kernel_histories)
        self.kernel = torch.from_numpy(k).to(self.device)
        self.net = ResidualCNN(seed=seed, device=self.device)
        self.bound = float(bound)
        self.max_exceed_frac = float(max_exceed_frac)
        # the network corrects voxels with plain dose >= region_frac * max (below the 10 %
        # evaluation threshold, so that region is covered); the plain dose stands elsewhere
        self.region_frac = float(region_frac)
        self._plans = {}
        self._bufs = {}

    @classmethod
    def pretrained(cls, nuclide: str = "Y90", voxel_mm: float = 2.0, **kw) -> "LearnedCorrector":
        """Corrector with the weights shipped in ``learned/weights`` (trained by
        ``learned/train.py`` on this package's Monte Carlo; see the file's metadata)."""
        name = PRETRAINED.get((nuclide, float(voxel_mm)))
        if name is None:
            raise ValueError(f"no pretrained weights for {nuclide} at {voxel_mm} mm; "
                             f"available: {sorted(PRETRAINED)}")
        path = os.path.join(WEIGHTS_DIR, name)
        from safetensors import safe_open
        with safe_open(path, framework="pt") as f:
            file_meta = dict(f.metadata() or {})
        # the Monte Carlo the weights were trained against (files without the key: electrons
        # only), the grid of the point kernel of their plain dose (without the key: 13) and
        # the correction bound chosen on held-out phantoms (without the key: 0.9)
        meta = dict(bound=float(file_meta.get("bound", "0.9")), photons=file_meta.get("mc_photons", "0") == "1",
                    kernel_size=int(file_meta.get("kernel_size", "13")))
        meta.update(kw)
        c = cls(nuclide, voxel_mm=voxel_mm, **meta)
        c.weights_meta = c.net.load_weights(path)
        return c

    def _stream(self):
        return ctypes.c_void_p(nat.current_stream_handle(self.device))

    def plan(self, shape):
        key = tuple(shape)
        if key not in self._plans:
            p = ConvPlan(key, tuple(self.kernel.shape), mode="linear", algo="fft", device=self.device)
            p.set_kernel(self.kernel)
            self._plans[key] = p
        return self._plans[key]

    def plain(self, act, rho, out=None):
        return self.plan(act.shape).dose(act, density=rho, out=out)

    def mc_dose(self, act, rho, histories: int = 64, seed: int = 1):
        return self.mc.dose(act, rho, self.voxel_mm, histories, seed)

    def _buf(self, shape, name, dtype, cols=None):
        import torch
        key = (tuple(shape), name)
        if key not in self._bufs:
            n = int(np.prod(shape))
            self._bufs[key] = torch.empty((n, cols) if cols else tuple(shape), dtype=dtype, device=self.device)
        return self._bufs[key]

    def input_maxima(self, act, dp):
        """(max activity, max dp) as 1-element device tensors, both in one pass
        (pvd_cnn_input_maxima); a fresh buffer per call, so earlier results stay valid."""
        import torch
        m = torch.empty(nat.PVD_CNN_MAXIMA_FLOATS, dtype=torch.float32, device=self.device)
        nat.check(self.lib.pvd_cnn_input_maxima(ctypes.c_void_p(act.data_ptr()), ctypes.c_void_p(dp.data_ptr()),
                                                act.numel(), ctypes.c_void_p(m.data_ptr()), self._stream()),
                  "pvd_cnn_input_maxima")
        return m[0:1], m[1:2]

    def pack(self, act, rho, dp, scales=None):
        """BF16 network input; returns (x, max dp as a 1-element device tensor).
        The scales are read on the device: no host synchronisation.  ``scales``:
        (max activity, max dp) device tensors to use instead of this input's own (a
        slab of a larger volume is normalised by the whole volume's maxima)."""
        import torch
        x = self._buf(act.shape, "x", torch.bfloat16, 16)
        if scales is None:
            amax, dmax = self.input_maxima(act, dp)
        else:
            amax, dmax = scales
        nat.check(self.lib.pvd_cnn_pack_inputs(ctypes.c_void_p(act.data_ptr()), ctypes.c_void_p(rho.data_ptr()),
                                               ctypes.c_void_p(dp.data_ptr()), ctypes.c_void_p(amax.data_ptr()),
                                               ctypes.c_void_p(dmax.data_ptr()), ctypes.c_void_p(x.data_ptr()),
                                               act.numel(), self._stream()),
                  "pvd_cnn_pack_inputs")
        self._keep = (amax, dmax)          # alive until the next call, past the asynchronous kernel
        return x, dmax

    def train_step(self, act, rho, target, lr: float = 1e-3, mask_frac: float = 0.05):
        """One Adam step on one (activity, density, target dose) triple; returns
        [sum m e^2, sum m, sum dr] of the relative-error loss e = r - (target/D_p - 1)
        on the voxels with D_p >= mask_frac * max D_p."""
        import torch
        shape = tuple(act.shape)
        dp = self.plain(act, rho, out=self._buf(shape, "dp", act.dtype))
        x, dmax = self.pack(act, rho, dp)
        self.net.forward(x, shape)
        dpf = dp.reshape(-1)
        mask = (dpf >= mask_frac * dmax).float()
        ones = torch.ones_like(dpf)
        ratio = torch.where(mask > 0, target.reshape(-1) / torch.clamp(dpf, min=1e-30), ones)
        sums = self.net.loss_and_grad(ones, ratio, mask, shape)
        self.net.backward(x, shape)
        self.net.adam_step(lr)
        return sums

    def region_blocks(self, dp, dmax, frac: float):
        """Device list of the network blocks within its receptive field of dp >= frac * max."""
        import torch
        shape = tuple(dp.shape)
        nb = int(self.lib.pvd_cnn_region_block_count(*shape))
        key = (shape, "region")
        if key not in self._bufs:
            self._bufs[key] = (torch.empty(nb, dtype=torch.int32, device=self.device),
                               torch.empty(nb, dtype=torch.int32, device=self.device),
                               torch.empty(1, dtype=torch.int32, device=self.device))
        flags, blist, count = self._bufs[key]
        nat.check(self.lib.pvd_cnn_region_blocks(ctypes.c_void_p(dp.data_ptr()), *shape, ctypes.c_void_p(dmax.data_ptr()),
                                                 ctypes.c_float(frac), ctypes.c_void_p(flags.data_ptr()),
                                                 ctypes.c_void_p(blist.data_ptr()), ctypes.c_void_p(count.data_ptr()),
                                                 self._stream()), "pvd_cnn_region_blocks")
        return blist, count

    def network_tiled(self, act, rho, dp, tile_z: int = 128, region: bool = False):
        """The network's correction r (FP32, the volume's shape) computed slab by slab
        along z: slabs of E voxels (E a multiple of 32, at most tile_z and Nz) that
        overlap by 2 x 6 voxels, each keeping the outputs at least 6 voxels (the
        receptive field of the six 3^3 layers) from its cut faces, so every output equals
        the whole-volume network's.  Inputs are normalised by the whole volume's maxima.
        Runs volumes whose Nz is not a multiple of 32, and bounds the activation
        memory by the slab (16 + 5 x 32 BF16 channels per voxel)."""
        import torch
        shape = tuple(act.shape)
        nz = shape[2]
        if nz < 32:
            raise ValueError(f"the network needs at least 32 voxels along z (got {nz})")
        e = min(int(tile_z) // 32 * 32, nz // 32 * 32)
        if e < 32:
            raise ValueError("tile_z must be at least 32")
        halo = 6
        amax, dmax = self.input_maxima(act, dp)
        r = torch.empty(shape, dtype=torch.float32, device=self.device)
        for s0, s1, a, b in z_slabs(nz, e, halo):
            sl = (slice(None), slice(None), slice(s0, s1))
            xs, _ = self.pack(act[sl].contiguous(), rho[sl].contiguous(), dp[sl].contiguous(), scales=(amax, dmax))
            sshape = (shape[0], shape[1], s1 - s0)
            blocks = self.region_blocks(dp[sl].contiguous(), dmax, self.region_frac) if region else None
            rs = self.net.forward(xs, sshape, keep_activations=False, blocks=blocks).reshape(sshape)
            r[:, :, a:b] = rs[:, :, a - s0:b - s0]
        self._keep_scales = (amax, dmax)
        return r.reshape(-1), dmax

    def correct(self, act, rho, r_override=None, region: bool = False, tile_z=None) -> Correction:
        """CNN-corrected dose with fallback (decided on the device: no host
        synchronisation).  With ``region`` the network runs only on the blocks within
        its receptive field of the region dp >= region_frac * max dp, and the correction
        applies there (the plain dose elsewhere; exact on that region).  The default
        runs it on the whole volume: the >= 10 %-max region of the Monte Carlo dose is
        not known before the correction, and in lung the plain dose's maximum can be
        several times the Monte Carlo one, so a plain-dose threshold can miss part of it.  `r_override` (tests) replaces the network output,
        e.g. to force an out-of-bound correction.  `plain` and `r` of the result are
        views of buffers the next call on this shape overwrites."""
        import torch
        shape = tuple(act.shape)
        dp = self.plain(act, rho, out=self._buf(shape, "dp", act.dtype))
        apply_frac = self.region_frac if region else 0.0
        if r_override is None and (tile_z is not None or shape[2] % 32):
            ext = tile_z or slab_extent(shape[2], shape[0] * shape[1])
            r, dmax = self.network_tiled(act, rho, dp, ext, region)
        elif r_override is None:
            x, dmax = self.pack(act, rho, dp)
            blocks = self.region_blocks(dp, dmax, apply_frac) if region else None
            r = self.net.forward(x, shape, keep_activations=False, blocks=blocks)
        else:
            dmax = dp.max().reshape(1)
            r = r_override.reshape(-1).contiguous()
        out = torch.empty(shape, dtype=torch.float32, device=self.device)
        counts = torch.zeros(2, dtype=torch.int64, device=self.device)
        nat.check(self.lib.pvd_cnn_apply_correction(ctypes.c_void_p(dp.data_ptr()), ctypes.c_void_p(r.data_ptr()),
                                                    ctypes.c_float(self.bound), ctypes.c_void_p(dmax.data_ptr()),
                                                    ctypes.c_float(0.1), ctypes.c_float(apply_frac),
                                                    ctypes.c_void_p(out.data_ptr()),
                                                    ctypes.c_void_p(counts.data_ptr()), dp.numel(), self._stream()),
                  "pvd_cnn_apply_correction")
        # the whole-volume fallback is decided and applied on the device: no host sync
        nat.check(self.lib.pvd_cnn_apply_fallback(ctypes.c_void_p(dp.data_ptr()), ctypes.c_void_p(counts.data_ptr()),
                                                  ctypes.c_double(self.max_exceed_frac),
                                                  ctypes.c_void_p(out.data_ptr()), dp.numel(), self._stream()),
                  "pvd_cnn_apply_fallback")
        return Correction(out, dp, r.reshape(shape), counts, self.max_exceed_frac)


def slab_extent(nz: int, plane_voxels: int, halo: int = 6, max_bytes: float = 32 * 2 ** 30) -> int:
    """The z extent (a multiple of 32, at most nz) whose slabs (`z_slabs`) cover nz with
    the fewest computed planes, among those whose activations (16 + 5 x 32 BF16
    channels per voxel) stay within max_bytes; the smaller extent on a tie.  For a
    512 x 512 x 400 volume: 224 (two slabs, 448 planes; slabs of 128 compute 512)."""
    best = None
    for e in range(32, nz // 32 * 32 + 1, 32):
        if e <= 2 * halo:
            continue
        if e > 32 and plane_voxels * e * (16 + 5 * 32) * 2 > max_bytes:
            break
        planes = sum(s1 - s0 for s0, s1, _, _ in z_slabs(nz, e, halo))
        if best is None or planes < best[0]:
            best = (planes, e)
    if best is None:
        raise ValueError(f"the network needs at least 32 voxels along z (got {nz})")
    return best[1]


def z_slabs(nz: int, e: int, halo: int = 6):
    """Slabs [s0, s1) of extent e covering [0, nz) for the tiled network, with the range
    [a, b) of each whose outputs are kept: at least `halo` voxels from a cut face (a
    volume face needs no margin).  Consecutive kept ranges tile [0, nz) exactly."""
    if e > nz or e <= 2 * halo:
        raise ValueError("slab extent must lie in (2 * halo, nz]")
    out = []
    s0, a = 0, 0
    while True:
        if s0 + e > nz:
            s0 = nz - e
        s1 = s0 + e
        b = nz if s1 == nz else s1 - halo
        out.append((s0, s1, a, b))
        if b == nz:
            return out
        a, s0 = b, b - halo


def voxel_errors(dose, ref, region_frac: float = 0.1):
    """Relative voxel errors |D - D_ref| / D_ref on D_ref >= region_frac * max(D_ref)."""
    region = ref >= region_frac * ref.max()
    rel = (dose[region] - ref[region]).abs() / ref[region]
    return rel


def subregion_box(act, size: int = 64, centre=None):
    """A size^3 box (slices) centred on `centre` (default: the activity maximum), shifted
    inside the volume."""
    import torch
    shape = tuple(act.shape)
    if min(shape) < size:
        raise ValueError(f"volume {shape} is smaller than a {size}^3 box")
    c = np.unravel_index(int(torch.argmax(act).item()), shape) if centre is None else centre
    lo = [min(max(int(ci) - size // 2, 0), n - size) for ci, n in zip(c, shape)]
    return tuple(slice(l, l + size) for l in lo)


def evaluate_subregion(corr: LearnedCorrector, act, rho, box, histories: int = 4096, seed: int = 1,
                       tol: float = 0.02, tile_z: int = 128):
    """Plain and CNN-corrected dose against Monte Carlo on a sub-region of a large volume.

    The activity is kept only inside `box`; the plain dose and the correction run on
    the whole volume (the network in z slabs), and the Monte Carlo transports the box's
    activity through the box's density.  All activity lies in the box, so inside it
    both doses have the same sources (particles that leave the box and scatter back are
    the only difference).  Errors on the box voxels >= 10 % of the Monte Carlo maximum."""
    import torch
    masked = torch.zeros_like(act)
    masked[box] = act[box]
    c = corr.correct(masked, rho, tile_z=tile_z)
    ref = corr.mc_dose(masked[box].contiguous(), rho[box].contiguous(), histories, seed=seed)
    plain, dose = c.plain[box], c.dose[box]
    rp, rc = voxel_errors(plain, ref), voxel_errors(dose, ref)
    out = {"plain_mean": float(rp.mean().item()), "cnn_mean": float(rc.mean().item()),
           "plain_p90": float(torch.quantile(rp, 0.9).item()), "cnn_p90": float(torch.quantile(rc, 0.9).item()),
           "plain_within": float((rp <= tol).float().mean().item()),
           "cnn_within": float((rc <= tol).float().mean().item()),
           "voxels": int(rp.numel()), "fallback": bool(c.fallback), "exceed": int(c.exceed),
           "box": [[b.start, b.stop] for b in box], "histories": int(histories)}
    return out


def evaluate(corr: LearnedCorrector, seeds, shape=(128, 128, 128), histories: int = 128, tol: float = 0.02):
    """Plain and CNN-corrected dose vs Monte Carlo on held-out phantoms."""
    import torch
    stats = {"plain_within": [], "cnn_within": [], "plain_mean": [], "cnn_mean": [], "plain_p90": [], "cnn_p90": [],
             "fallbacks": 0}
    for s in seeds:
        act, rho = synth_phantom(s, shape, corr.device)
        ref = corr.mc_dose(act, rho, histories, seed=s)
        c = corr.correct(act, rho)
        rp = voxel_errors(c.plain, ref)
        rc = voxel_errors(c.dose, ref)
        stats["plain_within"].append(float((rp <= tol).float().mean().item()))
        stats["cnn_within"].append(float((rc <= tol).float().mean().item()))
        stats["plain_mean"].append(float(rp.mean().item()))
        stats["cnn_mean"].append(float(rc.mean().item()))
        stats["plain_p90"].append(float(torch.quantile(rp, 0.9).item()))
        stats["cnn_p90"].append(float(torch.quantile(rc, 0.9).item()))
        stats["fallbacks"] += int(c.fallback)
    out = {k: (float(np.mean(v)) if isinstance(v, list) else v) for k, v in stats.items()}
    out["n"] = len(list(seeds)) if not isinstance(seeds, range) else len(seeds)
    torch.cuda.synchronize()
    return out


def time_per_volume(corr: LearnedCorrector, shape=(128, 128, 128), reps: int = 10):
    """GPU ms per volume of the plain dose and of the full CNN-corrected dose."""
    import torch
    act, rho = synth_phantom(123, shape, corr.device)
    for _ in range(3):
        corr.plain(act, rho)
        corr.correct(act, rho)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        corr.plain(act, rho)
    torch.cuda.synchronize()
    t_plain = (time.perf_counter() - t0) * 1e3 / reps
    x, _ = corr.pack(act, rho, corr.plain(act, rho))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        corr.net.forward(x, tuple(shape), keep_activations=False)
    torch.cuda.synchronize()
    t_net = (time.perf_counter() - t0) * 1e3 / reps
    t0 = time.perf_counter()
    for _ in range(reps):
        corr.correct(act, rho)
    torch.cuda.synchronize()
    t_corr = (time.perf_counter() - t0) * 1e3 / reps
    n = int(np.prod(shape))
    flops = 2.0 * 27 * (16 * 32 + 4 * 32 * 32 + 32 * 32) * n       # MFMA work as executed (padded channels)
    return {"plain_ms": t_plain, "cnn_forward_ms": t_net, "corrected_ms": t_corr,
            "cnn_tflops": flops / (t_net * 1e-3) / 1e12, "mfma_frac": flops / (t_net * 1e-3) / 2.5e15,
            "ratio_corrected_over_plain": t_corr / t_plain}
