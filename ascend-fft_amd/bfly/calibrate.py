"""Calibration of the plan-selection cost model (the role of the reference's
scripts/calib_eta.py:1-245, for the gfx950 kernels).

Model, per kernel key (include/butterfly/enumerate.hpp, KernelModel):

    eta_us(n, batch) = c0 + c1 * MB + c2 * Gpl

with MB = 16 n batch / 1e6 bytes moved (twice that for the four-step chain)
and Gpl = n batch log2(n) / 1e9 butterfly work.  c is the least-squares fit
over a measured sweep of the kernel; `loo` is the mean relative error of
leave-one-out predictions (each point predicted by the fit of the others).

    python -m bfly.calibrate --measure   # GPU: time the sweep, write
                                         # config/eta_sweep.json and the fit
                                         # config/cost_model.json (which
                                         # butterfly_space.json names)
    python -m bfly.calibrate --refit     # CPU: refit the committed sweep,
                                         # print it, --check: compare with the
                                         # committed coefficients
"""
from __future__ import annotations

import argparse
import json
import math
import sys
from pathlib import Path
from typing import Dict, List

import numpy as np

PKG = Path(__file__).resolve().parents[1]
if str(PKG) not in sys.path:
    sys.path.insert(0, str(PKG))

from bfly import _lib  # noqa: E402

SWEEP_FILE = PKG / "config" / "eta_sweep.json"
SPACE_FILE = PKG / "config" / "butterfly_space.json"
MODEL_FILE = PKG / "config" / "cost_model.json"    # named by the space file's Q.cost_model
LOCAL_MAX, SINGLE_MAX = 4096, 16384
KERNELS = {"plain": _lib.VARIANT_PLAIN, "stream": _lib.VARIANT_STREAM, "plain_swz": _lib.VARIANT_PLAIN_SWZ,
           "stream_swz": _lib.VARIANT_STREAM_SWZ, "stream_lt": _lib.VARIANT_STREAM_LT,
           "stream2": _lib.VARIANT_STREAM2, "four_step": _lib.VARIANT_FOUR_STEP,
           "split_mfma": _lib.VARIANT_SPLIT_MFMA, "split_mfma_stream": _lib.VARIANT_SPLIT_MFMA_STREAM}


def model_key(kernel: str, n: int) -> str:
    """The C++ modelKey(): one-workgroup forms of 8192 / 16384 get `_single`."""
    single = LOCAL_MAX < n <= SINGLE_MAX and kernel in ("plain", "stream_lt")
    return kernel + "_single" if single else kernel


def features(kernel: str, n: int, batch: int) -> List[float]:
    chain = kernel == "four_step" or n > SINGLE_MAX
    return [1.0, 16.0 * n * batch * (2.0 if chain else 1.0) / 1e6, n * batch * math.log2(n) / 1e9]


def sweep_points() -> List[dict]:
    """(kernel, n, batch) of the calibration sweep: each kernel over the
    lengths it has and four work sizes from launch-bound to HBM-bound."""
    pts = []
    totals = (1 << 16, 1 << 19, 1 << 22, 1 << 24)
    for kernel in ("plain", "stream", "plain_swz", "stream_swz", "stream_lt", "stream2"):
        for n in (64, 256, 1024, 4096):
            pts += [dict(kernel=kernel, n=n, batch=t // n) for t in totals]
    for kernel in ("split_mfma", "split_mfma_stream"):
        for n in (1024, 2048, 4096):
            pts += [dict(kernel=kernel, n=n, batch=t // n) for t in totals]
    for kernel in ("plain", "stream_lt"):
        for n in (8192, 16384):
            pts += [dict(kernel=kernel, n=n, batch=t // n) for t in (1 << 19, 1 << 21, 1 << 22, 1 << 24)]
    for n in (8192, 16384, 32768, 65536):
        pts += [dict(kernel="four_step", n=n, batch=t // n) for t in (1 << 19, 1 << 22, 1 << 24)]
    return pts


def fit(points: List[dict]) -> Dict[str, dict]:
    """Least-squares coefficients and leave-one-out error per model key."""
    groups: Dict[str, List[dict]] = {}
    for p in points:
        groups.setdefault(model_key(p["kernel"], p["n"]), []).append(p)
    out = {}
    for key, ps in sorted(groups.items()):
        X = np.array([features(p["kernel"], p["n"], p["batch"]) for p in ps])
        y = np.array([p["us"] for p in ps])
        c = np.linalg.lstsq(X, y, rcond=None)[0]
        errs = []
        for i in range(len(ps)):
            keep = np.arange(len(ps)) != i
            ci = np.linalg.lstsq(X[keep], y[keep], rcond=None)[0]
            errs.append(abs(float(X[i] @ ci) - y[i]) / y[i])
        fitted = X @ c
        out[key] = {"c": [round(float(v), 9) for v in c], "loo": round(float(np.mean(errs)), 6),
                    "loo_max": round(float(np.max(errs)), 6),
                    "fit_max": round(float(np.max(np.abs(fitted - y) / y)), 6), "points": len(ps)}
    return out


def measure(points: List[dict], reps: int = 20, rounds: int = 3, say=print) -> List[dict]:
    """Time every sweep point on the GPU: HIP events on the current stream,
    best of `rounds` means of `reps` launches that cycle through up to 16
    input / output pairs (at most 1 GiB of inputs, so the large points read
    from HBM; the small ones stay cache-resident as they would in use)."""
    import torch
    from bfly import ops
    lib = _lib.hip()
    dev = torch.device("cuda:0")
    s = torch.cuda.current_stream(dev)
    out = []
    for i, p in enumerate(points):
        kernel, n, b = p["kernel"], p["n"], p["batch"]
        v = KERNELS[kernel]
        floats = 2 * n * b
        pairs = max(2, min(16, (1 << 30) // (8 * floats)))
        xs = [torch.empty(floats, dtype=torch.float32, device=dev).uniform_(-1, 1) for _ in range(pairs)]
        ys = [torch.empty(floats, dtype=torch.float32, device=dev) for _ in range(pairs)]
        tab = ops.table(_lib.KIND_C2C, n, dev)
        work = None
        if lib.bfly_c2c_launches(n, b, v) == 2:
            work = torch.empty(floats, dtype=torch.float32, device=dev)

        def go(j):
            rc = lib.bfly_c2c_ex(xs[j % pairs].data_ptr(), ys[j % pairs].data_ptr(), n, b, -1, v, tab.data_ptr(),
                                 work.data_ptr() if work is not None else None, s.cuda_stream)
            _lib.check(rc, f"bfly_c2c_ex({kernel}, n={n}, batch={b})")
        for j in range(3):
            go(j)
        best = math.inf
        for _ in range(rounds):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(s)
            for j in range(reps):
                go(j)
            e1.record(s)
            e1.synchronize()
            best = min(best, e0.elapsed_time(e1) * 1e3 / reps)
        out.append(dict(p, us=round(best, 3)))
        say(f"[{i + 1}/{len(points)}] {kernel:18} n={n:<6} batch={b:<8} {best:9.2f} us")
        del xs, ys, work
    return out


def write_fit(models: Dict[str, dict], model_file: Path = MODEL_FILE) -> None:
    doc = {"model": "eta_us = c[0] + c[1] * MB moved + c[2] * n batch log2(n) / 1e9 per kernel key; least "
                    "squares over eta_sweep.json (python -m bfly.calibrate); loo = mean leave-one-out relative error",
           "kernels": {k: {"c": m["c"], "loo": m["loo"], "points": m["points"]} for k, m in models.items()}}
    model_file.write_text(json.dumps(doc, indent=1) + "\n")


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(prog="python -m bfly.calibrate", description=__doc__,
                                 formatter_class=argparse.RawDescriptionHelpFormatter)
    mode = ap.add_mutually_exclusive_group(required=True)
    mode.add_argument("--measure", action="store_true", help="time the sweep on the GPU and write the fit")
    mode.add_argument("--refit", action="store_true", help="refit the committed sweep (CPU)")
    ap.add_argument("--check", action="store_true", help="with --refit: fail unless the committed fit matches")
    ap.add_argument("--sweep", default=str(SWEEP_FILE))
    ap.add_argument("--model", default=str(MODEL_FILE))
    args = ap.parse_args(argv)
    if args.measure:
        pts = measure(sweep_points(), say=lambda t: print(t, flush=True))
        import torch
        Path(args.sweep).write_text(json.dumps({
            "device": torch.cuda.get_device_name(0), "timing": "HIP events, best of 3 means of 20 launches "
            "cycling through up to 16 buffer pairs (at most 1 GiB of inputs)", "points": pts}, indent=1) + "\n")
        models = fit(pts)
        write_fit(models, Path(args.model))
    else:
        models = fit(json.loads(Path(args.sweep).read_text())["points"])
    for k, m in models.items():
        print(f"{k:22} c={m['c']}  loo={m['loo']:.3f} (max {m['loo_max']:.3f})  fit max {m['fit_max']:.3f}  "
              f"points={m['points']}")
    if args.check:
        committed = json.loads(Path(args.model).read_text())["kernels"]
        bad = [k for k, m in models.items()
               if k not in committed or not np.allclose(committed[k]["c"], m["c"], rtol=1e-6, atol=1e-9)]
        if bad or set(committed) != set(models):
            print(f"committed cost model differs from the refit: {bad or sorted(set(committed) ^ set(models))}")
            return 1
    return 0


if __name__ == "__main__":
    sys.exit(main())
