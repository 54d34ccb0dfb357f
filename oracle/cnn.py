"""FP64 NumPy oracle of the 3-D residual CNN (test infrastructure only).

The network the GPU kernels implement (channels-last, 3x3x3 'same' zero padding,
cross-correlation ``y[v][co] = sum_{tap,ci} x[v + d(tap)][ci] w[tap][co][ci]``,
``tap = (dx+1)*9 + (dy+1)*3 + (dz+1)``):

    h0 = relu(conv0(x) + b0)
    h_l = h_{l-1} + relu(conv_l(h_{l-1}) + b_l),   l = 1..4
    r  = conv5(h4)[..., 0] + b5[0]

Loss of the multiplicative correction ``D = Dp (1 + r)`` against ``Dt`` on
voxel weights ``m``: ``L = sum m (Dp r - (Dt - Dp))^2 / sum m``.

Pinned by tests/test_oracle_cnn.py: ``conv3`` against ``scipy.ndimage.correlate``
per channel pair; ``backward`` against central finite differences of ``forward``;
``forward_stored`` with an identity store against ``forward``.
"""
from __future__ import annotations

import numpy as np

TAPS = [(dx, dy, dz) for dx in (-1, 0, 1) for dy in (-1, 0, 1) for dz in (-1, 0, 1)]


def bf16_round(a: np.ndarray) -> np.ndarray:
    """Round FP32 values to the nearest BF16 (ties to even), returned as FP32."""
    a32 = np.ascontiguousarray(a, dtype=np.float32)
    u = a32.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).reshape(a32.shape)


def conv3(x: np.ndarray, w: np.ndarray) -> np.ndarray:
    """x [n0,n1,n2,cin], w [27,cout,cin] -> [n0,n1,n2,cout] (FP64)."""
    x = np.asarray(x, np.float64)
    w = np.asarray(w, np.float64)
    n0, n1, n2, _ = x.shape
    xp = np.pad(x, ((1, 1), (1, 1), (1, 1), (0, 0)))
    y = np.zeros((n0, n1, n2, w.shape[1]))
    for t, (dx, dy, dz) in enumerate(TAPS):
        y += xp[1 + dx:1 + dx + n0, 1 + dy:1 + dy + n1, 1 + dz:1 + dz + n2, :] @ w[t].T
    return y


def conv3_input_grad(dy: np.ndarray, w: np.ndarray) -> np.ndarray:
    """dL/dx of conv3 given dL/dy [n0,n1,n2,cout]."""
    dy = np.asarray(dy, np.float64)
    n0, n1, n2, _ = dy.shape
    dyp = np.pad(dy, ((1, 1), (1, 1), (1, 1), (0, 0)))
    dx = np.zeros((n0, n1, n2, w.shape[2]))
    for t, (ax, ay, az) in enumerate(TAPS):
        # y[v] uses x[v + d]  =>  dx[u] += dy[u - d] w[t]
        dx += dyp[1 - ax:1 - ax + n0, 1 - ay:1 - ay + n1, 1 - az:1 - az + n2, :] @ np.asarray(w[t], np.float64)
    return dx


def conv3_weight_grad(x: np.ndarray, dy: np.ndarray) -> np.ndarray:
    """dL/dw [27,cout,cin] of conv3 given its input x and dL/dy."""
    x = np.asarray(x, np.float64)
    dy = np.asarray(dy, np.float64)
    n0, n1, n2, _ = x.shape
    xp = np.pad(x, ((1, 1), (1, 1), (1, 1), (0, 0)))
    dw = np.zeros((27, dy.shape[3], x.shape[3]))
    dyf = dy.reshape(-1, dy.shape[3])
    for t, (dx, dyy, dz) in enumerate(TAPS):
        xs = xp[1 + dx:1 + dx + n0, 1 + dyy:1 + dyy + n1, 1 + dz:1 + dz + n2, :].reshape(-1, x.shape[3])
        dw[t] = dyf.T @ xs
    return dw


def forward(params, x):
    """params: list of (w [27,32,cin], b [32]) for 6 layers. Returns (r, cache)."""
    hs = []
    acts = []
    z = conv3(x, params[0][0]) + params[0][1]
    a = np.maximum(z, 0.0)
    h = a
    hs.append(x)
    acts.append(a)
    for l in range(1, 5):
        hs.append(h)
        z = conv3(h, params[l][0]) + params[l][1]
        a = np.maximum(z, 0.0)
        acts.append(a)
        h = h + a
    hs.append(h)
    r = conv3(h, params[5][0])[..., 0] + params[5][1][0]
    return r, (hs, acts)


def forward_stored(params, x, store=bf16_round):
    """``forward`` with every activation the GPU stores between layers passed through
    ``store`` (BF16 rounding by default): h0 = store(relu(z0)), h_l = store(h_{l-1} +
    relu(z_l)); the sums themselves are FP64.  With ``store`` the identity it is
    ``forward`` (pinned in tests/test_oracle_cnn.py)."""
    h = store(np.maximum(conv3(x, params[0][0]) + params[0][1], 0.0))
    for l in range(1, 5):
        h = store(h + np.maximum(conv3(h, params[l][0]) + params[l][1], 0.0))
    return conv3(h, params[5][0])[..., 0] + params[5][1][0]


def loss_and_grad_r(r, dp, dt, m):
    e = dp * r - (dt - dp)
    cnt = m.sum()
    loss = (m * e * e).sum() / cnt
    return loss, 2.0 * m * e * dp / cnt


def backward(params, cache, dr):
    """Gradients [(dw, db)] for every layer given dL/dr."""
    hs, acts = cache
    grads = [None] * 6
    dy5 = np.zeros(dr.shape + (params[5][0].shape[1],))
    dy5[..., 0] = dr
    grads[5] = (conv3_weight_grad(hs[5], dy5), dy5.reshape(-1, dy5.shape[-1]).sum(0))
    dh = conv3_input_grad(dy5, params[5][0])
    for l in range(4, 0, -1):
        dz = dh * (acts[l] > 0)
        grads[l] = (conv3_weight_grad(hs[l], dz), dz.reshape(-1, 32).sum(0))
        dh = dh + conv3_input_grad(dz, params[l][0])
    dz = dh * (acts[0] > 0)
    grads[0] = (conv3_weight_grad(hs[0], dz), dz.reshape(-1, 32).sum(0))
    return grads


def init_params(rng, cin0=16, in_used=3, scale=1.0):
    """He-style initialisation; layer 0 only reads the first `in_used` input channels,
    layer 5 only produces channel 0 (rows co >= 1 are zero)."""
    params = []
    for l in range(6):
        cin = cin0 if l == 0 else 32
        fan = 27 * (in_used if l == 0 else 32)
        w = rng.normal(0.0, scale * np.sqrt(2.0 / fan), size=(27, 32, cin))
        if l == 0:
            w[:, :, in_used:] = 0.0
        b = np.zeros(32)
        if l == 5:
            w[:, 1:, :] = 0.0
            w *= 0.1
        params.append((w, b))
    return params
