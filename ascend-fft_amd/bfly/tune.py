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


VARIANT_NAMES = {
    _lib.VARIANT_PLAIN: "plain",
    _lib.VARIANT_STREAM: "stream",
    _lib.VARIANT_PLAIN_SWZ: "plain_swizzle",
    _lib.VARIANT_STREAM_SWZ: "stream_swizzle",
    _lib.VARIANT_STREAM_LT: "stream_lds_twiddles",
    _lib.VARIANT_STREAM2: "stream_depth2",
    _lib.VARIANT_SPLIT_MFMA: "split_fp16_mfma",
    _lib.VARIANT_SPLIT_MFMA_STREAM: "split_fp16_mfma_stream",
}


def variants(kind: int, n: int) -> list[int]:
    """C2C kernel variants of length n; the first is the plain one-shot grid."""
    if kind != _lib.KIND_C2C or n > 4096:
        return [_lib.VARIANT_PLAIN]
    mfma = (_lib.VARIANT_SPLIT_MFMA, _lib.VARIANT_SPLIT_MFMA_STREAM)
    return [v for v in VARIANT_NAMES if v not in mfma or n in (1024, 2048, 4096)]


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


def launch_variant(x, y, n: int, batch: int, variant: int, sign: int = -1) -> None:
    import torch
    from . import ops
    tab = ops.table(_lib.KIND_C2C, n, x.device)
    rc = _lib.hip().bfly_c2c_ex(x.data_ptr(), y.data_ptr(), n, batch, sign, variant, tab.data_ptr(),
                                None, torch.cuda.current_stream(x.device).cuda_stream)
    _lib.check(rc, f"bfly_c2c_ex(n={n}, batch={batch}, variant={variant})")


def time_variant(xs, ys, n: int, batch: int, variant: int, reps: int = 30) -> float:
    """Mean ms per launch over `reps` launches cycling through the buffer pairs."""
    import torch
    for i in range(3):
        launch_variant(xs[i % len(xs)], ys[i % len(ys)], n, batch, variant)
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    for i in range(reps):
        launch_variant(xs[i % len(xs)], ys[i % len(ys)], n, batch, variant)
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps


def chosen_vs_plain(n: int, batch: int, pairs, host_input: np.ndarray, sample_rows) -> dict:
    """Time every C2C variant of (n, batch) on the same rotated buffer pairs,
    gate each against FP64 on sample rows, and report the fastest passing one
    next to the plain one-shot variant."""
    import torch
    xs = [p[0] for p in pairs]
    ys = [p[1] for p in pairs]
    idx = torch.as_tensor(sample_rows, device=xs[0].device)
    results = []
    for v in variants(_lib.KIND_C2C, n):
        ms = time_variant(xs, ys, n, batch, v)
        launch_variant(xs[0], ys[0], n, batch, v)
        got = ys[0][idx].cpu().numpy()
        rel = rel_l2_fp64(got, host_input[sample_rows], n)
        results.append({"variant": v, "ms": ms, "rel_l2": rel})
    plain = results[0]
    ok = [r for r in results if r["rel_l2"] <= REL_L2_GATE]
    best = min(ok, key=lambda r: r["ms"]) if ok else plain
    cands = []
    for r in results:
        c = {"variant": VARIANT_NAMES.get(r["variant"], str(r["variant"])), "us": round(r["ms"] * 1e3, 2),
             "rel_l2": r["rel_l2"]}
        if r["variant"] in (_lib.VARIANT_SPLIT_MFMA, _lib.VARIANT_SPLIT_MFMA_STREAM):
            flop = split_mfma_flop(n, batch)
            c["mfma_tflops"] = round(flop / (r["ms"] * 1e-3) / 1e12, 1)
            c["mfma_util"] = round(flop / (r["ms"] * 1e-3) / FP16_DENSE_PEAK, 4)
        cands.append(c)
    return {
        "plan": VARIANT_NAMES.get(best["variant"], str(best["variant"])),
        "speedup_vs_plain": round(plain["ms"] / best["ms"], 4),
        "rel_l2_vs_fp64": best["rel_l2"],
        "candidates": cands,
    }


FP16_DENSE_PEAK = 2.5e15  # MI355X dense fp16 matrix flop/s (published, no sparsity)


def split_mfma_flop(n: int, batch: int) -> float:
    """Matrix flop of the split-FP16 path: per radix-16 pass, n/16 dense 16 x
    16 complex products per transform (a 32 x 32 real block times a 32-vector:
    2048 flop), three split products each."""
    passes = int(np.log2(n)) // 4
    return float(passes * (n // 16) * 2048 * 3 * batch)
