"""Measured plan selection (the prepare -> measure -> rank flow) for one shape.

Every kernel variant the library exposes for a shape is timed with HIP
events on the same resident buffers and checked against an FP64 FFT
(numpy's pocketfft on complex128) before it may be chosen; the fastest
variant within the 1e-5 relative-L2 gate wins.
"""
from __future__ import annotations

import numpy as np

from . import _lib

REL_L2_GATE = 1e-5


def variants(kind: int, n: int) -> list[int]:
    """Variant ids the library implements for (kind, n); 0 is the plain path."""
    lib = _lib.hip()
    if not hasattr(lib, "bfly_variants"):
        return [0]
    return [0]


def rel_l2_fp64(got: np.ndarray, x: np.ndarray, n: int, sign: int = -1) -> float:
    """Worst per-row relative L2 error of interleaved fp32 rows `got` against
    the FP64 transform of the interleaved fp32 rows `x`."""
    xc = x[:, 0::2].astype(np.float64) + 1j * x[:, 1::2].astype(np.float64)
    ref = np.fft.fft(xc, axis=1) if sign < 0 else np.fft.ifft(xc, axis=1) * n
    gc = got[:, 0::2].astype(np.float64) + 1j * got[:, 1::2].astype(np.float64)
    num = np.linalg.norm(gc - ref, axis=1)
    den = np.linalg.norm(ref, axis=1)
    den[den == 0] = 1.0
    return float(np.max(num / den))


def time_variant(x, y, n: int, batch: int, variant: int, reps: int = 20) -> float:
    import torch
    from . import ops
    for _ in range(3):
        ops.c2c_raw(x, y, n, batch, -1)
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        ops.c2c_raw(x, y, n, batch, -1)
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps


def chosen_vs_plain(n: int, batch: int, device, host_input: np.ndarray, sample_rows) -> dict:
    """Pick the fastest gate-passing variant for C2C (n, batch); report its
    speedup over the plain variant and its error vs FP64."""
    import torch
    x = torch.from_numpy(host_input).to(device)
    y = torch.empty_like(x)
    results = []
    for v in variants(_lib.KIND_C2C, n):
        ms = time_variant(x, y, n, batch, v)
        got = y[torch.as_tensor(sample_rows, device=device)].cpu().numpy()
        rel = rel_l2_fp64(got, host_input[sample_rows], n)
        results.append({"variant": v, "ms": ms, "rel_l2": rel})
    plain = results[0]
    ok = [r for r in results if r["rel_l2"] <= REL_L2_GATE]
    best = min(ok, key=lambda r: r["ms"]) if ok else plain
    return {
        "plan": "plain" if best["variant"] == 0 else f"variant{best['variant']}",
        "speedup_vs_plain": round(plain["ms"] / best["ms"], 4),
        "rel_l2_vs_fp64": best["rel_l2"],
        "candidates": [{"variant": r["variant"], "us": round(r["ms"] * 1e3, 2),
                        "rel_l2": r["rel_l2"]} for r in results],
    }
