"""Host <-> device array plumbing for the drop-in API.

The reference takes and returns NumPy arrays; this package also takes torch
tensors on the GPU and then returns tensors (no copies).
"""
from __future__ import annotations

import numpy as np


def to_device_f32(a, device=None):
    """Return (float32 contiguous tensor on the GPU, kind) with kind 'numpy'|'torch'."""
    import torch
    if isinstance(a, torch.Tensor):
        kind = "torch" if a.is_cuda else "numpy"
        dev = device if device is not None else (a.device if a.is_cuda else torch.device("cuda", torch.cuda.current_device()))
        t = a.to(device=dev, dtype=torch.float32)
        return t.contiguous(), kind
    arr = np.asarray(a)
    if arr.dtype.kind not in "fiub":
        raise TypeError(f"activity/density arrays must be real numeric, got {arr.dtype}")
    dev = device if device is not None else torch.device("cuda", torch.cuda.current_device())
    t = torch.from_numpy(np.ascontiguousarray(arr, dtype=np.float32)).to(dev)
    return t, "numpy"


def from_device(t, kind: str, dtype=np.float32):
    if kind == "torch":
        return t
    out = t.cpu().numpy()
    return out if out.dtype == dtype else out.astype(dtype)
