"""Measured choice of algorithm, FFT lengths, tile shapes and radix per (grid, kernel, mode).

Candidates: the FFT pipeline at the plan's default lengths (7-smooth, padded toward
lengths with a register-resident transform) and at the smallest 7-smooth lengths
that do not alias, and the direct LDS convolution when the
kernel extent allows.  For the fastest FFT lengths the tuner then walks the
tile size of each pass (8 / 16 / 32 lines per workgroup, z then y then x), the
largest Stockham radix (16 or 8) and the LDS Stockham stages in place of the
register-resident transforms, keeping a change only when it is at least 2 %
faster.  Each candidate is timed with HIP events on synthetic input;
the winner is cached in a JSON file keyed by shape, kernel, mode and device.
"""
from __future__ import annotations

import json
import os
import threading

from .plan import ConvPlan, describe_plan

_CACHE_PATH = os.environ.get("PVD_AUTOTUNE_CACHE",
                             os.path.join(os.path.expanduser("~"), ".cache", "pyvoxeldosimetry_amd", "autotune.json"))
_lock = threading.Lock()


def _smooth(n: int) -> bool:
    for p in (2, 3, 5, 7):
        while n % p == 0:
            n //= p
    return n == 1


def candidate_lengths(n: int, k: int, mode: str, count: int = 3):
    """Non-aliasing 7-smooth FFT lengths for one axis, smallest first."""
    if mode == "reference":
        ku = min(k, n)
        need = n + ku - 1
        out = [n] if _smooth(n) else []
    else:
        c = k // 2
        mlo, mhi = max(0, c - n + 1), min(k - 1, c + n - 1)
        need = n + max(mhi - c, c - mlo)
        out = []
    m = need
    while len(out) < count and m <= 2048:
        if _smooth(m):
            out.append(m)
        m += 1
    return out


def _load_cache():
    try:
        with open(_CACHE_PATH) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _save_cache(c):
    try:
        os.makedirs(os.path.dirname(_CACHE_PATH), exist_ok=True)
        tmp = _CACHE_PATH + ".tmp"
        with open(tmp, "w") as f:
            json.dump(c, f, indent=1, sort_keys=True)
        os.replace(tmp, _CACHE_PATH)
    except OSError:
        pass


def time_plan(plan: ConvPlan, reps: int = 10, warmup: int = 3) -> float:
    """Mean GPU ms per dose call on random input."""
    import torch
    a = torch.rand(plan.shape, dtype=torch.float32, device=plan.device)
    out = torch.empty_like(a)
    for _ in range(warmup):
        plan.dose(a, out=out)
    s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    s.record()
    for _ in range(reps):
        plan.dose(a, out=out)
    e.record()
    e.synchronize()
    return s.elapsed_time(e) / reps


def tuned_plan(shape, kernel_shape, mode, kernel_dev, max_candidates: int = 6, search_tiles: bool = True) -> ConvPlan:
    import torch
    dev = kernel_dev.device
    key = json.dumps([list(shape), list(kernel_shape), mode, torch.cuda.get_device_name(dev)])
    with _lock:
        cache = _load_cache()
        hit = cache.get(key)
    if hit is not None:
        fft = hit["algo"] == "fft"
        plan = ConvPlan(shape, kernel_shape, mode=mode, algo=hit["algo"], fft_n=hit.get("fft_n") if fft else None,
                        tile=hit.get("tile") if fft else None, max_radix=hit.get("max_radix", 0) if fft else 0,
                        stockham=bool(hit.get("stockham", False)) and fft, device=dev)
        plan.set_kernel(kernel_dev)
        return plan
    # the plan's own lengths (7-smooth, padded toward register-resident transforms)
    # first, then the smallest 7-smooth lengths per axis
    default = describe_plan(shape, kernel_shape, mode=mode, algo="fft")["fft_n"]
    cands = [("fft", tuple(default))]
    per_axis = [candidate_lengths(shape[i], kernel_shape[i], mode) for i in range(3)]
    seen = {tuple(default)}
    for i in range(max(len(a) for a in per_axis)):
        fn = tuple(a[min(i, len(a) - 1)] for a in per_axis)
        if fn not in seen:
            seen.add(fn)
            cands.append(("fft", fn))
    cands = cands[:max_candidates]
    if all(min(shape[i], kernel_shape[i]) <= 16 if mode == "reference" else kernel_shape[i] <= 16 for i in range(3)):
        cands.append(("direct", None))
    best = None
    for algo, fn in cands:
        try:
            plan = ConvPlan(shape, kernel_shape, mode=mode, algo=algo, fft_n=fn, device=dev)
        except (ValueError, NotImplementedError, MemoryError):
            continue
        plan.set_kernel(kernel_dev)
        t = time_plan(plan)
        if best is None or t < best[0]:
            if best is not None:
                best[1].close()
            best = (t, plan, algo, fn)
        else:
            plan.close()
    if best is None:
        raise RuntimeError(f"no usable convolution plan for shape {shape}, kernel {kernel_shape}")
    tile, radix, stockham = list(best[1].tile) if best[2] == "fft" else None, 16, False
    if best[2] == "fft" and search_tiles:
        fn = best[1].fft_n

        def trial(t, r, lds_only=False):
            try:
                p = ConvPlan(shape, kernel_shape, mode=mode, algo="fft", fft_n=fn, tile=t, max_radix=r,
                             stockham=lds_only, device=dev)
            except (ValueError, NotImplementedError, MemoryError):
                return None
            p.set_kernel(kernel_dev)
            return time_plan(p), p

        # from the tiles the plan chose itself (16 lines, 8 on x for some two-step lengths)
        for axis in (2, 1, 0):
            for t in (8, 16, 32):
                if t == tile[axis]:
                    continue
                cand = list(tile)
                cand[axis] = t
                res = trial(cand, radix)
                if res is not None and res[0] < 0.98 * best[0]:
                    best[1].close()
                    best = (res[0], res[1], "fft", fn)
                    tile = cand
                elif res is not None:
                    res[1].close()
        res = trial(tile, 8)
        if res is not None and res[0] < 0.98 * best[0]:
            best[1].close()
            best = (res[0], res[1], "fft", fn)
            radix = 8
        elif res is not None:
            res[1].close()
        res = trial(tile, radix, True)
        if res is not None and res[0] < 0.98 * best[0]:
            best[1].close()
            best = (res[0], res[1], "fft", fn)
            stockham = True
        elif res is not None:
            res[1].close()
    with _lock:
        cache = _load_cache()
        cache[key] = {"algo": best[2], "fft_n": list(best[3]) if best[3] else None, "ms": best[0],
                      "tile": tile if best[2] == "fft" else None, "max_radix": radix if best[2] == "fft" else 0,
                      "stockham": stockham}
        _save_cache(cache)
    return best[1]
