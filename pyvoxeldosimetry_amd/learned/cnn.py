"""3-D residual CNN for the heterogeneity correction, on the HIP kernels of
``include/pvd_learned.h`` (BF16 MFMA, FP32 accumulation, FP32 master weights).

    h0 = relu(conv0(x) + b0)                 x: 16 channels (activity, density, plain dose, 13 zeros)
    h_l = h_{l-1} + relu(conv_l(h_{l-1}) + b_l),   l = 1..4   (32 channels)
    r  = conv5(h4)[..., 0] + b5[0]            relative correction: D = D_plain (1 + r)

Six 3x3x3 layers, 32 channels; 114,209 trainable parameters (layer 0 reads
3 input channels, layer 5 writes one).
"""
from __future__ import annotations

import ctypes
import math

from .. import _native as nat

CIN = [16, 32, 32, 32, 32, 32]
IN_USED = 3


def _p(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else ctypes.c_void_p(0)


class ResidualCNN:
    def __init__(self, seed: int = 0, device=None):
        import torch
        nat.require_device()
        self.lib = nat.load()
        self.device = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
        self.offsets = []
        off = 0
        for cin in CIN:
            self.offsets.append((off, off + 27 * 32 * cin))      # weight slice, then 32 biases
            off += 27 * 32 * cin + 32
        self.n_params = off
        g = torch.Generator().manual_seed(seed)
        master = torch.zeros(off, dtype=torch.float32)
        for l, cin in enumerate(CIN):
            w0, w1 = self.offsets[l]
            fan = 27 * (IN_USED if l == 0 else 32)
            w = torch.randn((27, 32, cin), generator=g) * math.sqrt(2.0 / fan)
            if l == 0:
                w[:, :, IN_USED:] = 0.0
            if l == 5:
                w.zero_()          # the correction starts at r = 0: the plain dose
            master[w0:w1] = w.reshape(-1)
        self.master = master.to(self.device)
        self.bf16 = self.master.to(torch.bfloat16)
        self.grad = torch.zeros_like(self.master)
        self.m = torch.zeros_like(self.master)
        self.v = torch.zeros_like(self.master)
        self.step = 0
        self._ws = {}
        self._flip = torch.empty(27 * 32 * 32, dtype=torch.bfloat16, device=self.device)

    # ------------------------------------------------------------------ parameters
    def trainable_parameters(self) -> int:
        """Parameters that can be non-zero (layer 0 reads 3 channels, layer 5 writes 1)."""
        n = 27 * 32 * IN_USED + 32 + 4 * (27 * 32 * 32 + 32) + 27 * 32 + 1
        return n

    def w_bf16(self, l):
        a, b = self.offsets[l]
        return self.bf16[a:b]

    def bias(self, l):
        _, b = self.offsets[l]
        return self.master[b:b + 32]

    def grad_w(self, l):
        a, b = self.offsets[l]
        return self.grad[a:b]

    def grad_b(self, l):
        _, b = self.offsets[l]
        return self.grad[b:b + 32]

    def state_dict(self):
        return {"master": self.master.detach().cpu(), "m": self.m.cpu(), "v": self.v.cpu(), "step": self.step}

    def load_state_dict(self, sd):
        import torch
        self.master.copy_(sd["master"].to(self.device))
        self.bf16.copy_(self.master.to(torch.bfloat16))
        self.m.copy_(sd["m"].to(self.device))
        self.v.copy_(sd["v"].to(self.device))
        self.step = int(sd["step"])

    def save_weights(self, path: str, metadata=None) -> None:
        """FP32 master weights only, as safetensors (data, no code)."""
        from safetensors.torch import save_file
        meta = {k: str(v) for k, v in (metadata or {}).items()}
        meta["layout"] = "flat: per layer w[27][32][cin] then b[32]; cin = " + ",".join(map(str, CIN))
        save_file({"master": self.master.detach().cpu().contiguous()}, path, metadata=meta)

    def load_weights(self, path: str) -> dict:
        """Load weights written by save_weights; resets the optimiser state. Returns the metadata."""
        import torch
        from safetensors import safe_open
        with safe_open(path, framework="pt") as f:
            meta = dict(f.metadata() or {})
            master = f.get_tensor("master")
        if master.dtype != torch.float32 or master.numel() != self.n_params:
            raise ValueError(f"{path}: expected {self.n_params} FP32 parameters, got {master.numel()} {master.dtype}")
        self.master.copy_(master.to(self.device))
        self.bf16.copy_(self.master.to(torch.bfloat16))
        self.m.zero_()
        self.v.zero_()
        self.step = 0
        return meta

    # ------------------------------------------------------------------ workspace
    def _workspace(self, shape):
        import torch
        key = tuple(shape)
        ws = self._ws.get(key)
        if ws is None:
            n = shape[0] * shape[1] * shape[2]
            bf = dict(dtype=torch.bfloat16, device=self.device)
            ws = {
                # zeroed once: a region-restricted forward reads neighbours it did not write
                "h": [torch.zeros((n, 32), **bf) for _ in range(5)],
                "act": [None] + [torch.empty((n, 32), **bf) for _ in range(4)],
                "r": torch.zeros(n, dtype=torch.float32, device=self.device),
                "dz": torch.empty((n, 32), **bf),
                "dr": torch.empty((n, 32), **bf),
                "dh": torch.empty((n, 32), dtype=torch.float32, device=self.device),
                "dbias_part": torch.empty(((n + 1023) // 1024, 32), dtype=torch.float32, device=self.device),
                "wwork": torch.empty(int(self.lib.pvd_cnn_wgrad_work_floats(*shape)), dtype=torch.float32,
                                     device=self.device),
                "loss_part": torch.empty((256, 3), dtype=torch.float32, device=self.device),
                "loss_sum": torch.empty(3, dtype=torch.float32, device=self.device),
            }
            self._ws[key] = ws
        return ws

    def _stream(self):
        return ctypes.c_void_p(nat.current_stream_handle(self.device))

    def _conv(self, x, cin, l, epi, shape, y=None, act=None, resid=None, y32=None, weights=None, bias=True,
              blocks=None):
        bl, bn = blocks if blocks is not None else (None, None)
        nat.check(self.lib.pvd_cnn_conv_fwd(_p(x), cin, _p(weights if weights is not None else self.w_bf16(l)),
                                            _p(self.bias(l)) if bias else ctypes.c_void_p(0), epi, _p(resid),
                                            _p(y), _p(act), _p(y32), shape[0], shape[1], shape[2], _p(bl), _p(bn),
                                            self._stream()),
                  "pvd_cnn_conv_fwd")

    # ------------------------------------------------------------------ forward / backward
    def forward(self, x, shape, keep_activations: bool = True, blocks=None):
        """x: BF16 [n, 16] inputs; returns the FP32 correction r [n] (workspace tensor).
        ``keep_activations=False`` (inference) skips the ReLU outputs backward needs.
        ``blocks``: (block list, count) device tensors from ``pvd_cnn_region_blocks``:
        every layer runs on those blocks only, and r is defined on the region they cover."""
        ws = self._workspace(shape)
        h = ws["h"]
        # the input's channels 4..15 are zero (pvd_cnn_pack): the four-channel input layer
        self._conv(x, 4, 0, nat.PVD_CNN_EPI_RELU, shape, y=h[0], blocks=blocks)
        for l in range(1, 5):
            self._conv(h[l - 1], 32, l, nat.PVD_CNN_EPI_RESID_RELU, shape, y=h[l],
                       act=ws["act"][l] if keep_activations else None, resid=h[l - 1], blocks=blocks)
        self._conv(h[4], 32, 5, nat.PVD_CNN_EPI_LINEAR_F32, shape, y32=ws["r"], blocks=blocks)
        return ws["r"]

    def loss_and_grad(self, dp, dt, mask, shape):
        """Loss on the last forward; writes dL/dr into the workspace. dp, dt, mask: FP32 [n]."""
        import torch
        ws = self._workspace(shape)
        n = ws["r"].numel()
        cnt = float(mask.sum().item())
        if cnt <= 0:
            raise ValueError("empty loss mask")
        nat.check(self.lib.pvd_cnn_loss(_p(ws["r"]), _p(dp), _p(dt), _p(mask), n, ctypes.c_float(1.0 / cnt),
                                        _p(ws["dr"]), _p(ws["loss_part"]), self._stream()), "pvd_cnn_loss")
        nat.check(self.lib.pvd_reduce_rows(_p(ws["loss_part"]), 256, 3, _p(ws["loss_sum"]), 0, self._stream()),
                  "pvd_reduce_rows")
        return ws["loss_sum"]        # [sum m e^2, sum m, sum dr]; loss = [0] / [1]

    def backward(self, x, shape):
        """Gradients of every parameter into self.grad (after forward + loss_and_grad)."""
        ws = self._workspace(shape)
        n = ws["r"].numel()
        h, act, dz, dh = ws["h"], ws["act"], ws["dz"], ws["dh"]
        st = self._stream()
        # layer 5
        nat.check(self.lib.pvd_cnn_wgrad(_p(h[4]), 32, _p(ws["dr"]), _p(self.grad_w(5)), _p(ws["wwork"]),
                                         shape[0], shape[1], shape[2], st), "pvd_cnn_wgrad")
        gb5 = self.grad_b(5)
        gb5.zero_()
        gb5[0:1].copy_(ws["loss_sum"][2:3])
        dh.zero_()
        nat.check(self.lib.pvd_cnn_flip_weights(_p(self.w_bf16(5)), _p(self._flip), st), "pvd_cnn_flip_weights")
        self._conv(ws["dr"], 32, 5, nat.PVD_CNN_EPI_ACCUM_F32, shape, y32=dh, weights=self._flip, bias=False)
        for l in range(4, -1, -1):
            a = act[l] if l > 0 else h[0]
            nat.check(self.lib.pvd_cnn_relu_grad(_p(dh), _p(a), _p(dz), _p(ws["dbias_part"]), n, st),
                      "pvd_cnn_relu_grad")
            nat.check(self.lib.pvd_reduce_rows(_p(ws["dbias_part"]), ws["dbias_part"].shape[0], 32,
                                               _p(self.grad_b(l)), 0, st), "pvd_reduce_rows")
            xin = x if l == 0 else h[l - 1]
            nat.check(self.lib.pvd_cnn_wgrad(_p(xin), CIN[l], _p(dz), _p(self.grad_w(l)), _p(ws["wwork"]),
                                             shape[0], shape[1], shape[2], st), "pvd_cnn_wgrad")
            if l > 0:
                nat.check(self.lib.pvd_cnn_flip_weights(_p(self.w_bf16(l)), _p(self._flip), st),
                          "pvd_cnn_flip_weights")
                self._conv(dz, 32, l, nat.PVD_CNN_EPI_ACCUM_F32, shape, y32=dh, weights=self._flip, bias=False)
        # parameters that are structurally zero keep zero gradients
        self.grad_w(0).view(27, 32, 16)[:, :, IN_USED:] = 0.0
        self.grad_w(5).view(27, 32, 32)[:, 1:, :] = 0.0

    def adam_step(self, lr: float = 1e-3, beta1: float = 0.9, beta2: float = 0.999, eps: float = 1e-8):
        self.step += 1
        nat.check(self.lib.pvd_cnn_adam(_p(self.master), _p(self.grad), _p(self.m), _p(self.v), _p(self.bf16),
                                        self.n_params, ctypes.c_float(lr), ctypes.c_float(beta1),
                                        ctypes.c_float(beta2), ctypes.c_float(eps), self.step, self._stream()),
                  "pvd_cnn_adam")

    def params_numpy(self):
        """[(w [27,32,cin], b [32])] as FP64 NumPy, weights as the BF16 values used on the GPU."""
        out = []
        for l, cin in enumerate(CIN):
            w = self.w_bf16(l).float().cpu().double().numpy().reshape(27, 32, cin)
            b = self.bias(l).cpu().double().numpy()
            out.append((w, b))
        return out
