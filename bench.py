#!/usr/bin/env python3
"""Headline benchmark: plain fp32 C2C forward, N=4096, batch=4096, on one MI355X.

One step = one bfly_c2c launch (the library's automatic variant choice) over
the whole batch ([4096][8192] float32, interleaved), inputs resident in HBM.  Prints ONE JSON line:
  value          algorithmic GB/s = 16*N*batch bytes / mean step time
  roofline       achieved GB/s of the kernel (HIP events on its stream) vs the
                 8 TB/s HBM peak, plus HBM traffic per launch from the committed
                 rocprofv3 counter summary (profiles/), or null
  cpu_baseline   the reference's own FP64 CPU FFT (oracle/_ref, built from the
                 reference sources) timed on this host on a bounded sample
  chosen         every kernel variant of this shape timed on the same buffers
                 and gated against FP64; the fastest passing one, its speedup
                 over the plain one-shot variant and its relative error
  rows           the other hot-path rows at their shapes: C2C n=1024 b=64
                 latency, R2C / C2R 4096x4096 (reference CPU rows beside
                 them), long C2C 65536 (b=64, 256) and 16384 (b=1024), padded-row
                 C2C and the normalised inverse at 4096 x 4096
  patterns       the uneven batched workload (512 requests) as one grouped
                 persistent launch under hipGraph vs per-request launches:
                 transforms/s, GB/s, p50/p99 per-request latency; the
                 spectral-convolution layer (fwd+bwd) vs the same layer on
                 torch.fft; the host-to-host 4096 x 4096 call over PCIe
                 (staged vs pipelined, beside the reference's end-to-end row)

Timed steps rotate over several input/output buffer pairs (>= 1 GiB in
total) so that the 256 MB last-level cache cannot serve one step's input from
the previous step.
"""
from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "ascend-fft_amd"))

N = 4096
BATCH = 4096
HBM_PEAK_GBS = 8000.0
METRIC = "C2C fp32 GB/s (%HBM peak) N=4096 b=4096; chosen/plain speedup @rel-err 1e-5"


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--rotate", type=int, default=4, help="input/output buffer pairs cycled")
    ap.add_argument("--prewarm-seconds", type=float, default=0.5,
                    help="untimed steps before the warmup steps (GPU clock ramp)")
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="CPU baseline budget")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--only", default=None,
                    choices=["grouped", "r2c", "c2r", "mfma", "long65536", "single16384", "spectral", "fft2d"],
                    help="run one secondary workload loop instead of the headline (profiling)")
    ap.add_argument("--no-patterns", action="store_true",
                    help="skip the grouped uneven-batch measurement")
    ap.add_argument("--probe-copy", action="store_true",
                    help="also run the streaming copy kernel on 256 MiB (counter calibration)")
    return ap.parse_args()


def traffic_from_profile():
    """HBM bytes per launch from the committed rocprofv3 counter summary."""
    p = ROOT / "profiles" / "c2c_n4096_b4096_pmc.md"
    if not p.exists():
        return None
    m = re.search(r"traffic_bytes_per_launch:\s*([0-9.eE+]+)", p.read_text())
    return float(m.group(1)) if m else None


def cpu_baseline(budget_s: float):
    """The reference's refFftF32 (FP64 radix-2, one core) over a bounded
    sample of the same workload: passes over the 4096 seeded rows until the
    time budget is spent."""
    import numpy as np
    import oracle
    r = oracle.ref()
    kind = "reference"
    if r is not None:
        fn = r.ref_fft_f32
    else:
        fn = oracle.lib().orc_ref_fft_f32
        kind = "port"
    x = oracle.uniform_input(0, (BATCH, 2 * N))
    rows_in = [np.ascontiguousarray(x[i]) for i in range(BATCH)]
    out = np.empty(2 * N, np.float32)
    t0 = time.perf_counter()
    done = 0
    while time.perf_counter() - t0 < budget_s:
        for i in range(256):
            fn(rows_in[(done + i) % BATCH], out, N)
        done += 256
    dt = time.perf_counter() - t0
    return {
        "value": round(16.0 * N * done / dt / 1e9, 4),
        "unit": "GB/s",
        "cores": 1,
        "kind": kind,
        "sample": f"{done} transforms ({done / BATCH:.2f} passes over the 4096 seeded rows) of "
                  f"the N=4096 C2C workload through the reference's refFftF32 (FP64, one "
                  f"thread) in {dt:.1f} s",
    }


def cpu_libraries(budget_s: float = 2.0):
    """Host FFT libraries on the headline workload (n = 4096, batch = 4096,
    fp32 complex rows: 128 MiB, far above the host caches), each for a
    bounded time: torch.fft.fft on CPU tensors with every host thread,
    scipy.fft with every host thread, numpy.fft on one thread.  Calls take
    512-row blocks in order through the whole tensor, so no block is reused
    before the other 112 MiB were touched.  Context for the reference's own
    CPU figure (cpu_baseline), not a target."""
    import os

    import numpy as np
    import scipy.fft
    import torch

    cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    rows = 512
    x = np.random.default_rng(0).uniform(-1, 1, (BATCH, 2 * N)).astype(np.float32).view(np.complex64)
    blocks = [x[i:i + rows] for i in range(0, BATCH, rows)]
    tblocks = [torch.from_numpy(b) for b in blocks]
    old = torch.get_num_threads()
    torch.set_num_threads(cores)
    runs = {"torch.fft.fft (CPU)": (lambda i: torch.fft.fft(tblocks[i % len(tblocks)]), cores),
            "scipy.fft.fft": (lambda i: scipy.fft.fft(blocks[i % len(blocks)], workers=cores), cores),
            "numpy.fft.fft": (lambda i: np.fft.fft(blocks[i % len(blocks)]), 1)}
    out = {}
    try:
        for name, (fn, thr) in runs.items():
            fn(0)
            t0, done = time.perf_counter(), 0
            while time.perf_counter() - t0 < budget_s:
                fn(done // rows + 1)
                done += rows
            dt = time.perf_counter() - t0
            out[name] = {"GB_per_s": round(16.0 * N * done / dt / 1e9, 3), "threads": thr,
                         "sample": f"{done} rows of n={N} in {dt:.1f} s, 512-row blocks taken in order "
                                   f"through the full {BATCH}-row (128 MiB) input"}
    finally:
        torch.set_num_threads(old)
    return out


def clock_up(fn, seconds: float = 0.25, per_sync: int = 20):
    """Untimed: run fn(i) back to back for `seconds` of wall time.  The GPU
    leaves its idle clocks only after some milliseconds of sustained work, so
    a workload timed right after host-only work (the CPU baselines) with a
    few warm-up launches reads up to ~12 % slow (C2R n=4096 b=4096: 31.4 us
    after a pause against 27.7 us clocked up)."""
    import torch
    t0, i = time.perf_counter(), 0
    while time.perf_counter() - t0 < seconds:
        for _ in range(per_sync):
            fn(i)
            i += 1
        torch.cuda.synchronize()


def copy_ceiling():
    """The device's measured streaming ceiling, as the hardware profile
    records it: bfly_probe_copy (the non-temporal four-deep float4 copy) at
    256 MiB and 1 GiB each way, best of three rounds of twenty copies per
    size (HIP events), the faster size (bfly.probe.copy_bandwidth)."""
    from bfly import probe
    c = probe.copy_bandwidth()
    return c["gbs"], c["best_mib"]


def secondary_rows(cpu_seconds: float, reps: int = 20, rounds: int = 3):
    """The other hot-path rows at the shapes the reference quotes, inputs resident and
    rotated over 4 buffer sets (>= 256 MB MALL): cfg1 C2C latency, R2C / C2R
    4096x4096 (with the reference's refR2CF32 / refC2RF32 timed on one host
    core over a bounded sample), long C2C at 65536 (b = 64 and 256) and
    16384 (b = 1024).  Kernel time by HIP events on the launch stream, the
    reference's protocol (docs/benchmarks/methodology.md:52-55): warm-up
    (0.25 s of back-to-back launches, clock_up), then `rounds` rounds of `reps` launches, the minimum of the round means."""
    import numpy as np
    import torch
    from bfly import _lib, ops
    import oracle

    lib = _lib.hip()
    dev = torch.device("cuda:0")
    sid = lambda: torch.cuda.current_stream().cuda_stream  # noqa: E731

    def timed(fn):
        clock_up(fn)
        best = None
        for _ in range(rounds):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            for i in range(reps):
                fn(i)
            e1.record()
            e1.synchronize()
            ms = e0.elapsed_time(e1) / reps
            best = ms if best is None else min(best, ms)
        return best

    def row(ms, nbytes, transforms):
        gbs = nbytes / (ms * 1e-3) / 1e9
        return {"us": round(ms * 1e3, 2), "GB_per_s": round(gbs, 1), "frac_of_8TBs": round(gbs / HBM_PEAK_GBS, 4),
                "transforms_per_s": round(transforms / (ms * 1e-3), 1), "algorithmic_bytes": int(nbytes)}

    out = {}
    # cfg1: C2C n=1024 b=64 (latency-bound): back-to-back launches from C
    # (bfly_rt_time_transform, the Plan::measure loop), so Python call
    # overhead does not set the rate
    import ctypes
    n, b = 1024, 64
    x1 = torch.rand(b, 2 * n, device=dev)
    y1 = torch.empty_like(x1)
    tab1 = ops.table(_lib.KIND_C2C, n, dev)
    ms1 = ctypes.c_float()
    for _ in range(2):
        _lib.check(lib.bfly_rt_time_transform(_lib.KIND_C2C, -1, x1.data_ptr(), y1.data_ptr(), n, b,
                                              tab1.data_ptr(), None, 200, sid(), ctypes.byref(ms1)),
                   "bfly_rt_time_transform")
    out["c2c_n1024_b64"] = row(ms1.value / 200, 16 * n * b, b)
    # R2C / C2R 4096 x 4096
    n, b = 4096, 4096
    xr = [torch.rand(b, n, device=dev) * 2 - 1 for _ in range(4)]
    hs = [torch.empty(b, n + 2, device=dev) for _ in range(4)]
    for kind, name, src, dst in ((1, "r2c_n4096_b4096", xr, hs), (2, "c2r_n4096_b4096", hs, xr)):
        tab = ops.table(kind, n, dev)
        # arguments prepared once: the host loop must outpace the ~27 us kernels
        args = [(kind, src[i].data_ptr(), dst[i].data_ptr(), n, b, n // 2 + 1, n + 2, _lib.VARIANT_AUTO, 0,
                 tab.data_ptr(), sid()) for i in range(4)]
        out[name] = row(timed(lambda i: _lib.check(lib.bfly_real_ex(*args[i % 4]), "real")),
                        (4 * n + 4 * (n + 2)) * b, b)
    del xr, hs
    r = oracle.ref()
    if r is not None and cpu_seconds > 0:
        x = oracle.uniform_input(1, (64, n))
        h = oracle.r2c_f64(x, n).astype(np.float32)
        h[:, 1] = 0
        h[:, n + 1] = 0
        for name, fn, a, width in (("r2c_n4096_b4096", r.ref_r2c_f32, x, n + 2),
                                   ("c2r_n4096_b4096", r.ref_c2r_f32, h, n)):
            o = np.empty(width, np.float32)
            t0, done = time.perf_counter(), 0
            while time.perf_counter() - t0 < cpu_seconds:
                fn(np.ascontiguousarray(a[done % 64]), o, n)
                done += 1
            dt = time.perf_counter() - t0
            out[name]["cpu_baseline"] = {
                "value": round((4 * n + 4 * (n + 2)) * done / dt / 1e9, 4), "unit": "GB/s", "cores": 1,
                "kind": "reference", "sample": f"{done} rows of n={n} through the reference's "
                f"{'refR2CF32' if name.startswith('r2c') else 'refC2RF32'} in {dt:.1f} s"}
    # long C2C (library calls with their arguments prepared once)
    for n, b in ((65536, 64), (65536, 256), (16384, 1024)):
        xs = [torch.rand(b, 2 * n, device=dev) for _ in range(4)]
        ys = [torch.empty_like(xs[0]) for _ in range(4)]
        work = ops.workspace(0, n, b, dev)
        tab = ops.table(_lib.KIND_C2C, n, dev)
        args = [(xs[i].data_ptr(), ys[i].data_ptr(), n, b, -1, tab.data_ptr(),
                 work.data_ptr() if work is not None else None, sid()) for i in range(4)]
        out[f"c2c_n{n}_b{b}"] = row(timed(lambda i: _lib.check(lib.bfly_c2c(*args[i % 4]), "long")),
                                    16 * n * b, b)
        del xs, ys, work
    # padded rows (transforms 4096 + 64 complex values apart, gaps untouched)
    # and the normalised inverse (scale 1/n in the last store), both at the
    # headline shape
    n, b, d = 4096, 4096, 4096 + 64
    xs = [torch.rand(b, 2 * d, device=dev) for _ in range(4)]
    ys = [torch.empty_like(xs[0]) for _ in range(4)]
    tab = ops.table(_lib.KIND_C2C, n, dev)
    args = [(xs[i].data_ptr(), d, ys[i].data_ptr(), d, n, b, -1, tab.data_ptr(), None, sid()) for i in range(4)]
    out["c2c_rows_n4096_b4096_dist4160"] = row(
        timed(lambda i: _lib.check(lib.bfly_c2c_rows(*args[i % 4]), "rows")), 16 * n * b, b)
    args = [(xs[i].data_ptr(), ys[i].data_ptr(), n, b, 1, 1.0 / n, tab.data_ptr(), None, sid()) for i in range(4)]
    out["c2c_inv_normalised_n4096_b4096"] = row(
        timed(lambda i: _lib.check(lib.bfly_c2c_scaled(*args[i % 4]), "scaled")), 16 * n * b, b)
    del xs, ys
    return out


def grouped_pattern(reps: int = 20):
    """The uneven batched workload: 512 C2C/R2C requests (n log-uniform over
    64..8192, batch U[1,256], seed 3) as ONE grouped persistent launch
    captured in a hipGraph, configuration chosen by the autotuner, against
    the same requests as 512 per-request launches (eager and captured).
    Outputs are checked against the per-request launches before timing."""
    import dataclasses

    import numpy as np
    import torch
    from bfly import _lib, group

    specs = group.uneven_specs()
    reqs = group.make_requests(specs, seed=3)
    group.launch_per_request(reqs)
    torch.cuda.synchronize()
    ref = group.outputs(reqs)
    # serving objective: the lowest median request latency among the
    # configurations within 5 % of the fastest
    tuned = group.autotune(reqs, reps=5, reference=ref, latency_slack=0.05)
    fastest = min((row for row in tuned.table if row[2] <= 1e-5), key=lambda row: row[1])
    g = group.Group(reqs, tuned.best)
    for r in reqs:
        r.y.fill_(float("nan"))
    g.launch()
    torch.cuda.synchronize()
    err = group.max_rel_diff(group.outputs(reqs), ref)
    graph = group.capture(g.launch)
    ms_group = group.time_ms(graph.replay, reps=reps)
    # the per-request baseline: one library call per request, the library's
    # own choice of kernel (AUTO)
    auto = _lib.VARIANT_AUTO
    per_graph = group.capture(lambda: group.launch_per_request(reqs, auto))
    ms_per_graph = group.time_ms(per_graph.replay, reps=reps)
    ms_per_eager = group.time_ms(lambda: group.launch_per_request(reqs, auto), reps=max(3, reps // 4))
    lat = group.latencies_us(g)
    lat_seq = group.per_request_latencies_us(reqs, auto)
    # the same configuration with the least-work-first unit order (latency)
    cfg_lat = dataclasses.replace(tuned.best, order=3)
    g_lat = group.Group(reqs, cfg_lat)
    ms_lat = group.time_ms(g_lat.launch, reps=reps)
    lat_sjf = group.latencies_us(g_lat)
    g_lat.close()
    nbytes = g.algorithmic_bytes()
    transforms = g.transforms()
    kinds = [k for k, _, _ in specs]
    out = {
        "workload": "512 requests, n log-uniform 64..8192, batch U[1,256], C2C/R2C 50/50, seed 3",
        "requests": len(reqs),
        "c2c_requests": int(sum(1 for k in kinds if k == 0)),
        "transforms": transforms,
        "algorithmic_bytes": nbytes,
        "config": tuned.best.name(),
        "objective": "lowest median request latency within 5 % of the fastest gated configuration",
        "fastest_config": {"config": fastest[0], "ms": round(fastest[1], 4),
                           "latency_us_p50": None if fastest[5] is None else round(fastest[5], 2)},
        "units": g.units,
        "grid": g.grid,
        "ms_grouped_graph": round(ms_group, 4),
        "transforms_per_s": round(transforms / (ms_group * 1e-3), 1),
        "GB_per_s": round(nbytes / (ms_group * 1e-3) / 1e9, 1),
        "pct_hbm_peak": round(100.0 * nbytes / (ms_group * 1e-3) / 1e9 / HBM_PEAK_GBS, 2),
        "latency_us_p50": round(float(np.percentile(lat, 50)), 2),
        "latency_us_p99": round(float(np.percentile(lat, 99)), 2),
        "least_work_first": {
            "config": cfg_lat.name(),
            "ms": round(ms_lat, 4),
            "latency_us_p50": round(float(np.percentile(lat_sjf, 50)), 2),
            "latency_us_p99": round(float(np.percentile(lat_sjf, 99)), 2),
        },
        "per_request_launches": {
            "variant": "auto",
            "ms_graph": round(ms_per_graph, 4),
            "ms_eager": round(ms_per_eager, 4),
            "latency_us_p50": round(float(np.percentile(lat_seq, 50)), 2),
            "latency_us_p99": round(float(np.percentile(lat_seq, 99)), 2),
        },
        "speedup_vs_per_request_graph": round(ms_per_graph / ms_group, 3),
        "max_rel_diff_vs_per_request": err,
        "tuned_candidates": [{"config": n, "ms": round(ms, 4), "rel_diff": e,
                              "latency_us_p50": None if p50 is None else round(p50, 2)}
                             for n, ms, e, _, _, p50 in tuned.table],
    }
    del graph, per_graph
    g.close()
    return out


def spectral_pattern(reps: int = 10):
    """The 1-D spectral-convolution layer, x [64, 128, 4096] ~ N(0,1),
    W [128, 128, 256] complex ~ N(0, 1/256) per part (seed 4): forward +
    backward on the mode-truncated kernels vs the same layer on torch.fft
    (hipFFT), outputs and gradients compared in FP64."""
    import torch
    from bfly.spectral import spectral_conv1d, torch_fft_reference

    torch.manual_seed(4)
    B, C, N, K = 64, 128, 4096, 256
    x = torch.randn(B, C, N, device="cuda")
    s = (1.0 / (C * 2)) ** 0.5
    w = torch.complex(torch.randn(C, C, K, device="cuda") * s, torch.randn(C, C, K, device="cuda") * s)
    up = torch.randn(B, C, N, device="cuda")

    def run(layer):
        xa, wa = x.clone().requires_grad_(True), w.clone().requires_grad_(True)
        y = layer(xa, wa)
        (y * up).sum().backward()
        return y.detach(), xa.grad, wa.grad

    def timed(layer):
        for _ in range(2):
            run(layer)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            run(layer)
        e1.record()
        e1.synchronize()
        return e0.elapsed_time(e1) / reps

    ours, ref = run(spectral_conv1d), run(torch_fft_reference)
    rel = [float((a.double() - b.double()).norm() / b.double().norm()) for a, b in zip(ours, ref)]
    ms_ours, ms_ref = timed(spectral_conv1d), timed(torch_fft_reference)
    return {
        "workload": "spectral conv1d fwd+bwd, x [64,128,4096], W [128,128,256] complex, seed 4",
        "ms_fwd_bwd": round(ms_ours, 3),
        "ms_fwd_bwd_torch_fft": round(ms_ref, 3),
        "speedup_vs_torch_fft": round(ms_ref / ms_ours, 3),
        "rel_l2_vs_torch_fft": {"y": rel[0], "dx": rel[1], "dw": rel[2]},
    }


def grid49_pattern(reps: int = 50):
    """The plain drop-in grid of 49 points, C2C forward fp32, n in
    {64..4096} x batch in {1,4,...,4096} (inputs U[-1,1) from seed 1): the
    library (AUTO kernel) against torch.fft.fft (hipFFT) on the same GPU and
    the same device-resident complex64 tensors, both launched back to back
    from Python and timed with HIP events; geometric-mean speedup and wins, as
    the reference reports its grid against its vendor FFT."""
    import numpy as np
    import torch
    from bfly import _lib, ops

    lib = _lib.hip()
    dev = torch.device("cuda:0")
    sid = torch.cuda.current_stream().cuda_stream
    g = torch.Generator(device=dev)
    g.manual_seed(1)

    def timed(fn):
        for _ in range(5):
            fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        e1.synchronize()
        return e0.elapsed_time(e1) / reps * 1e3

    points, ratios = [], []
    for n in (64, 128, 256, 512, 1024, 2048, 4096):
        tab = ops.table(_lib.KIND_C2C, n, dev)
        for b in (1, 4, 16, 64, 256, 1024, 4096):
            x = torch.complex(torch.rand(b, n, generator=g, device=dev) * 2 - 1,
                              torch.rand(b, n, generator=g, device=dev) * 2 - 1)
            y = torch.empty_like(x)
            args = (x.data_ptr(), y.data_ptr(), n, b, -1, tab.data_ptr(), None, sid)
            ours = timed(lambda: lib.bfly_c2c(*args))
            ref = timed(lambda: torch.fft.fft(x))
            err = float((y - torch.fft.fft(x)).abs().max() / torch.fft.fft(x).abs().max())
            points.append({"n": n, "batch": b, "us": round(ours, 2), "us_torch_fft": round(ref, 2),
                           "speedup": round(ref / ours, 3), "max_rel_diff": err})
            ratios.append(ref / ours)
            del x, y
    return {
        "workload": "C2C fp32 grid n 64..4096 x batch 1..4096 (49 points), seed 1, vs torch.fft.fft (hipFFT)",
        "geomean_speedup_vs_torch_fft": round(float(np.exp(np.mean(np.log(ratios)))), 3),
        "wins": int(sum(r > 1.0 for r in ratios)),
        "losses": int(sum(r <= 1.0 for r in ratios)),
        "points": points,
    }


def fft2d_pattern(reps: int = 30):
    """2-D C2C fp32 on the reference's declared 2-D shapes (fft2d squares and
    range-Doppler, config/test_matrix.json:94-99), batch sized to 256 MiB per
    tensor: the library (bfly_fft2d) against torch.fft.fft2 (hipFFT) on the
    same device tensors, HIP events; algorithmic bytes 16 n1 n2 batch."""
    import torch
    from bfly import ops

    dev = torch.device("cuda:0")
    g = torch.Generator(device=dev)
    g.manual_seed(6)

    def timed(fn):
        for _ in range(3):
            fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        e1.synchronize()
        return e0.elapsed_time(e1) / reps * 1e3

    points = []
    for n1, n2 in ((256, 256), (1024, 1024), (2048, 2048), (1024, 128), (2048, 256), (4096, 256)):
        b = max(1, (32 << 20) // (n1 * n2))
        x = torch.complex(torch.rand(b, n1, n2, generator=g, device=dev) * 2 - 1,
                          torch.rand(b, n1, n2, generator=g, device=dev) * 2 - 1)
        y = torch.empty_like(x)
        work = torch.empty(2 * x.numel(), dtype=torch.float32, device=dev)
        xf, yf = torch.view_as_real(x), torch.view_as_real(y)
        ours = timed(lambda: ops.fft2_raw(xf, yf, n1, n2, b, -1, work))
        ref = timed(lambda: torch.fft.fft2(x))
        want = torch.fft.fft2(x)
        err = float(torch.linalg.vector_norm(y - want) / torch.linalg.vector_norm(want))
        nbytes = 16 * n1 * n2 * b
        points.append({"n1": n1, "n2": n2, "batch": b, "us": round(ours, 2), "GB_per_s": round(nbytes / ours / 1e3, 1),
                       "hbm_round_trips": 2 if n1 <= 2048 else 3, "us_torch_fft2": round(ref, 2),
                       "speedup": round(ref / ours, 3), "rel_l2_vs_torch": err})
        del x, y, work
    return {"workload": "2-D C2C fp32 [batch][n1][n2], 256 MiB per tensor, seed 6, vs torch.fft.fft2 (hipFFT)",
            "points": points}


def host_e2e_pattern(reps: int = 20):
    """C2C fp32 N=4096 b=4096 from pinned host memory to pinned host memory
    (the reference's end-to-end row: pinned H2D + FFT + D2H, 7306.3 us on its
    hardware): the staged upload -> kernel -> download, each step waited for,
    against bfly_host_pipelined (8 MiB chunks, uploads / transforms /
    downloads overlapped), on a stream of the library's own as the drop-in
    Plan::run uses; wall clock per call, mean / median / min over `reps`
    after three untimed calls; the output of the last pipelined call checked against the
    staged one bit for bit.  PCIe-inclusive, never the headline value."""
    import numpy as np
    import torch
    from bfly import _lib, ops

    dev = torch.device("cuda:0")
    lib = _lib.hip()
    n, b = N, BATCH
    nbytes = 8 * n * b
    hin = torch.empty(b, 2 * n, dtype=torch.float32, pin_memory=True)
    hin.copy_(torch.from_numpy(np.random.default_rng(7).uniform(-1, 1, (b, 2 * n)).astype(np.float32)))
    hout = torch.empty(b, 2 * n, dtype=torch.float32, pin_memory=True)
    hout2 = torch.empty(b, 2 * n, dtype=torch.float32, pin_memory=True)
    din = torch.empty(b * 2 * n, dtype=torch.float32, device=dev)
    dout = torch.empty_like(din)
    tab = ops.table(_lib.KIND_C2C, n, dev)
    # a stream of the library's own, as bfly::Context holds for Plan::run
    import ctypes
    sp = ctypes.c_void_p()
    _lib.check(lib.bfly_rt_stream_create(ctypes.byref(sp)), "bfly_rt_stream_create")
    s = sp.value
    variant = _lib.VARIANT_STREAM_LT

    def staged():
        _lib.check(lib.bfly_rt_memcpy_h2d(din.data_ptr(), hin.data_ptr(), nbytes, s), "h2d")
        _lib.check(lib.bfly_c2c_ex(din.data_ptr(), dout.data_ptr(), n, b, -1, variant, tab.data_ptr(), None, s),
                   "bfly_c2c_ex")
        _lib.check(lib.bfly_rt_stream_sync(s), "sync")
        _lib.check(lib.bfly_rt_memcpy_d2h(hout.data_ptr(), dout.data_ptr(), nbytes, s), "d2h")

    def pipelined():
        _lib.check(lib.bfly_host_pipelined(_lib.KIND_C2C, -1, variant, hin.data_ptr(), hout2.data_ptr(), n, b,
                                           tab.data_ptr(), din.data_ptr(), dout.data_ptr(), None, 0, s),
                   "bfly_host_pipelined")

    def timed(fn):
        for _ in range(3):  # the first calls set up queues and map the pinned pages
            fn()
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            fn()
            ts.append((time.perf_counter() - t0) * 1e6)
        return float(np.mean(ts)), float(np.median(ts)), float(np.min(ts))

    torch.cuda.synchronize()
    st_mean, st_med, st_min = timed(staged)
    pl_mean, pl_med, pl_min = timed(pipelined)
    _lib.check(lib.bfly_rt_stream_destroy(s), "bfly_rt_stream_destroy")
    same = bool(torch.equal(hout, hout2))
    ref_us = 7306.3
    return {"workload": "C2C fp32 N=4096 b=4096, pinned host in -> pinned host out (128 MiB each way), "
                        "wall clock per call",
            "staged_us": round(st_mean, 1), "staged_median_us": round(st_med, 1), "staged_min_us": round(st_min, 1),
            "pipelined_us": round(pl_mean, 1), "pipelined_median_us": round(pl_med, 1),
            "pipelined_min_us": round(pl_min, 1),
            "pipelined_GB_per_s_both_ways": round(2 * nbytes / pl_mean / 1e3, 1),
            "pipelined_equals_staged": same,
            "reference_e2e_us": ref_us, "reference_source": "published end-to-end row, Ascend910 (pinned H2D + "
                                                            "FFT + D2H)",
            "speedup_vs_reference": round(ref_us / pl_mean, 2)}


def workload_loop(name: str, steps: int, warmup: int):
    """One secondary workload run `steps` times after `warmup` untimed runs,
    for kernel traces and counter passes of its kernels (--only NAME).
    Prints one JSON line with the mean ms per step (HIP events)."""
    import numpy as np
    import torch
    from bfly import _lib, group, ops

    lib = _lib.hip()
    dev = torch.device("cuda:0")
    sid = lambda: torch.cuda.current_stream().cuda_stream  # noqa: E731
    g = torch.Generator(device=dev)
    g.manual_seed(5)

    def rand(*shape):
        return torch.rand(*shape, generator=g, device=dev) * 2 - 1

    # the single-kernel workloads rotate over 4 input/output sets (> the
    # 256 MB MALL), as the timed rows of the full run do
    rot = [0]

    def rotating(calls):
        def fn():
            calls[rot[0] % 4]()
            rot[0] += 1
        return fn

    if name == "spectral":
        from bfly.spectral import spectral_conv1d
        torch.manual_seed(4)
        B, Cc, Nn, K = 64, 128, 4096, 256
        xs = torch.randn(B, Cc, Nn, device=dev)
        sc = (1.0 / (Cc * 2)) ** 0.5
        ws = torch.complex(torch.randn(Cc, Cc, K, device=dev) * sc, torch.randn(Cc, Cc, K, device=dev) * sc)
        up = torch.randn(B, Cc, Nn, device=dev)

        def fn():
            xa, wa = xs.clone().requires_grad_(True), ws.clone().requires_grad_(True)
            (spectral_conv1d(xa, wa) * up).sum().backward()
        nbytes = 4 * (2 * B * Cc * Nn) + 8 * Cc * Cc * K
    elif name == "grouped":
        reqs = group.make_requests(group.uneven_specs(), seed=3)
        grp = group.Group(reqs, group.Config(order=3, dynamic=1, twiddles=1))
        graph = group.capture(grp.launch)
        fn, nbytes = graph.replay, grp.algorithmic_bytes()
    elif name == "fft2d":
        n1, n2, b = 1024, 1024, 16
        xs = [rand(b, n1, 2 * n2) for _ in range(4)]
        ys = [torch.empty(b, n1, 2 * n2, device=dev) for _ in range(4)]
        work = torch.empty(2 * b * n1 * n2, dtype=torch.float32, device=dev)
        fn = rotating([(lambda i: lambda: ops.fft2_raw(xs[i], ys[i], n1, n2, b, -1, work))(i) for i in range(4)])
        nbytes = 16 * n1 * n2 * b
    elif name in ("r2c", "c2r"):
        n, b = 4096, 4096
        xs = [rand(b, n) for _ in range(4)]
        hs = [rand(b, n + 2) for _ in range(4)]
        for h in hs:
            h[:, 1] = 0
            h[:, n + 1] = 0
        kind = 1 if name == "r2c" else 2
        tab = ops.table(kind, n, dev)
        src, dst = (xs, hs) if kind == 1 else (hs, xs)
        args = [(kind, src[i].data_ptr(), dst[i].data_ptr(), n, b, n // 2 + 1, n + 2, _lib.VARIANT_AUTO, 0,
                 tab.data_ptr(), sid()) for i in range(4)]
        fn = rotating([(lambda a: lambda: _lib.check(lib.bfly_real_ex(*a), name))(a) for a in args])
        nbytes = (4 * n + 4 * (n + 2)) * b
    elif name in ("mfma", "long65536", "single16384"):
        n, b, v = {"mfma": (4096, 4096, _lib.VARIANT_SPLIT_MFMA), "long65536": (65536, 256, _lib.VARIANT_AUTO),
                   "single16384": (16384, 1024, _lib.VARIANT_AUTO)}[name]
        xs = [rand(b, 2 * n) for _ in range(4)]
        ys = [torch.empty_like(xs[0]) for _ in range(4)]
        tab = ops.table(0, n, dev)
        work = ops.workspace(0, n, b, dev)
        wp = work.data_ptr() if work is not None else None
        args = [(xs[i].data_ptr(), ys[i].data_ptr(), n, b, -1, v, tab.data_ptr(), wp, sid()) for i in range(4)]
        fn = rotating([(lambda a: lambda: _lib.check(lib.bfly_c2c_ex(*a), name))(a) for a in args])
        nbytes = 16 * n * b
    else:
        raise SystemExit(f"unknown workload {name}")
    clock_up(lambda i: fn())
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(steps):
        fn()
    e1.record()
    e1.synchronize()
    ms = e0.elapsed_time(e1) / steps
    print(json.dumps({"workload": name, "steps": steps, "ms_per_step": round(ms, 5),
                      "algorithmic_bytes": int(nbytes), "GB_per_s": round(nbytes / (ms * 1e-3) / 1e9, 1)}))


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this benchmark runs on exactly one GPU; --gpus N>1 is not "
                                   "supported (batch rows would shard with no communication)"}))
        sys.exit(2)
    if args.only:
        workload_loop(args.only, args.steps, args.warmup)
        return
    import numpy as np
    import torch
    from bfly import ops, tune

    dev = torch.device("cuda:0")
    torch.manual_seed(0)
    pairs = []
    rng = np.random.default_rng(0)
    host = rng.uniform(-1.0, 1.0, size=(BATCH, 2 * N)).astype(np.float32)
    for i in range(max(1, args.rotate)):
        x = torch.from_numpy(host).to(dev) if i == 0 else torch.roll(pairs[0][0], shifts=i, dims=0)
        pairs.append((x.contiguous(), torch.empty_like(x)))
    stream = torch.cuda.current_stream(dev)
    # One step = one bfly_c2c call through the C ABI, its arguments prepared
    # once (as a C or FFI caller holds them): the host then spends a few us
    # per launch, so the timed region is the kernels plus one launch latency
    # and one synchronisation, not Python argument marshalling.
    from bfly import _lib
    lib = _lib.hip()
    tab = ops.table(_lib.KIND_C2C, N, dev)
    assert lib.bfly_workspace_bytes(_lib.KIND_C2C, N, BATCH) == 0
    step_args = [(x.data_ptr(), y.data_ptr(), N, BATCH, -1, tab.data_ptr(), None, stream.cuda_stream)
                 for x, y in pairs]
    c2c = lib.bfly_c2c

    def step(i):
        rc = c2c(*step_args[i % len(step_args)])
        if rc:
            _lib.check(rc, f"bfly_c2c(n={N}, batch={BATCH})")

    # untimed: bring the GPU out of its idle clocks before the warmup steps
    # (a short warmup alone leaves the first timed steps a few % slow)
    t_pre = time.perf_counter()
    i = 0
    while time.perf_counter() - t_pre < args.prewarm_seconds:
        for _ in range(50):
            step(i)
            i += 1
        torch.cuda.synchronize()
    for i in range(args.warmup):
        step(i)
    torch.cuda.synchronize()
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    ev0.record(stream)
    for i in range(args.steps):
        step(i)
    ev1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kernel_ms = ev0.elapsed_time(ev1) / args.steps
    ms_per_step = wall * 1e3 / args.steps
    bytes_alg = 16.0 * N * BATCH
    value = bytes_alg / (ms_per_step * 1e-3) / 1e9
    achieved = bytes_alg / (kernel_ms * 1e-3) / 1e9

    # accuracy of the plain path (outside the timed region)
    x0, y0 = pairs[0]
    ops.c2c_raw(x0, y0, N, BATCH, -1)
    torch.cuda.synchronize()
    sample = np.arange(0, BATCH, 64)
    got = y0[sample].cpu().numpy()
    rel = tune.rel_l2_fp64(got, host[sample], N)

    chosen = tune.chosen_vs_plain(N, BATCH, pairs, host, sample)
    ceiling, ceiling_mib = copy_ceiling()

    if args.probe_copy:
        # Known-size traffic for calibrating FETCH_SIZE / WRITE_SIZE:
        # 256 MiB read + 256 MiB written per launch.
        src = torch.empty(64 << 20, dtype=torch.float32, device=dev).uniform_()
        dst = torch.empty_like(src)
        for _ in range(3):
            ops.probe_copy(src, dst)
        torch.cuda.synchronize()

    line = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic: U[-1,1) interleaved complex, numpy default_rng(0), rotated over "
                f"{len(pairs)} resident buffer pairs",
        "config": {"workload": "c2c_fwd_fp32_n4096_b4096", "n": N, "batch": BATCH,
                   "layout": "interleaved [batch][2n] float32", "parallelism": "single GPU"},
        "pct_hbm_peak": round(100.0 * value / HBM_PEAK_GBS, 2),
        "transforms_per_s": round(BATCH / (ms_per_step * 1e-3), 1),
        "rel_l2_vs_fp64": rel,
        "chosen": chosen,
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic_from_profile(),
            "kernel_us": round(kernel_ms * 1e3, 3),
            "algorithmic_bytes_per_launch": int(bytes_alg),
            "measured_ceiling": {"value": ceiling, "unit": "GB/s",
                                 "source": f"bfly_probe_copy: non-temporal four-deep float4 copy, best of 256 MiB "
                                           f"and 1 GiB each way ({ceiling_mib} MiB), this run"},
            "frac_of_measured_ceiling": round(achieved / ceiling, 4),
        },
        "cpu_baseline": None if args.no_cpu else cpu_baseline(args.cpu_seconds),
    }
    if not args.no_cpu:
        line["cpu_libraries"] = cpu_libraries()
    if not args.no_patterns:
        line["rows"] = secondary_rows(0 if args.no_cpu else 1.5)
        line["patterns"] = {"grouped_uneven": grouped_pattern(), "spectral_conv1d": spectral_pattern(),
                            "grid49_vs_torch_fft": grid49_pattern(), "fft2d_vs_torch_fft2": fft2d_pattern(),
                            "host_end_to_end": host_e2e_pattern()}
    print(json.dumps(line))


if __name__ == "__main__":
    main()
