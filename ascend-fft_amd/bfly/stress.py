"""Budgeted stress runner for the `stress` profile of config/test_matrix.json.

The counterpart of the reference's scripts/run_stress.py:1-627, over the
bfly_check program and the C ABI:

* shapes come from the declared input-buffer tiers (batch = tier bytes /
  bytes of one input row), and each is budgeted before anything is
  allocated: the exact device bytes bfly_check allocates (input, output,
  table, four-step workspace) against the free device memory less a
  headroom, the host bytes (its buffers, the table, the input scratch)
  against MemAvailable less a margin, and the 32-bit float-offset guard;
  a refused row keeps its numbers and the reason;
* every tier runs one canonical C2C length as a soak of `repeat_execution`
  launches of one plan, sampling the process RSS: growth beyond
  max(96 MiB, 8 % of the peak) fails the soak;
* a lifecycle stage starts and ends the program `--lifecycle-cycles` times
  (device init, plan, run, teardown), each run checked;
* the opt-in `--oom-probe` asks the device for twice its memory through
  bfly_rt_malloc (refused at once: nothing is exhausted), then proves the
  same process still allocates and transforms correctly.

    python -m bfly.stress [--plan] [--soak-reps K] [--lifecycle-cycles C]
                          [--oom-probe --device D] [--out DIR] [--exe PATH]
"""
import argparse
import csv
import json
import math
import os
import re
import subprocess
import sys
import time
from datetime import datetime, timezone
from pathlib import Path

PKG = Path(__file__).resolve().parents[1]
if str(PKG) not in sys.path:
    sys.path.insert(0, str(PKG))

from bfly.matrix import DEFAULT_CONFIG, DEFAULT_EXE, command_for, load_config, verdict_of  # noqa: E402

MIB = 1 << 20
GIB = 1 << 30
OFFSET_LIMIT = 1 << 32
HOST_MARGIN = 0.02
DEVICE_HEADROOM = 0.05
QUICK_REPS = 3
RSS_SAMPLE_S = 0.02
RSS_MIN_SAMPLES = 8
FIELDS = ("stage", "tier_bytes", "direction", "n", "batch", "reps", "input",
          "input_bytes", "host_avail", "device_free", "device_total",
          "host_needed", "device_needed", "decision", "reason", "returncode",
          "wall_ms", "max_rel", "rss_growth_mb", "correct", "failure")


def input_bytes_for(direction, n, batch):
    """Bytes of the input tensor: interleaved complex, n reals, or n/2+1 bins."""
    if direction == "c2r":
        return 4 * (n + 2) * batch
    if direction == "r2c":
        return 4 * n * batch
    return 8 * n * batch


def output_bytes_for(direction, n, batch):
    if direction == "c2r":
        return 4 * n * batch
    if direction == "r2c":
        return 4 * (n + 2) * batch
    return 8 * n * batch


def derive_batch(n, tier_bytes, direction):
    if n < 1 or tier_bytes < 1:
        return 0
    return tier_bytes // input_bytes_for(direction, n, 1)


def _host_lib():
    """A library for the host-only shape queries (no device call): the gfx950
    one when it is built, else the CPU stub (same host code)."""
    from bfly import _lib
    try:
        return _lib.hip()
    except RuntimeError:
        return _lib.stub()


def table_bytes(direction, n):
    """Device table of one (kind, n): bfly_table_floats * 4."""
    from bfly import _lib
    kind = {"c2c": _lib.KIND_C2C, "r2c": _lib.KIND_R2C, "c2r": _lib.KIND_C2R}[direction]
    return 4 * int(_host_lib().bfly_table_floats(kind, n))


def four_step(n, batch):
    """True when bfly_check's C2C of (n, batch) runs the four-step chain (and
    so allocates the n*batch complex workspace)."""
    return _host_lib().bfly_c2c_launches(n, batch, 0) == 2


def device_bytes_exact(direction, n, batch):
    dev = input_bytes_for(direction, n, batch) + output_bytes_for(direction, n, batch) + table_bytes(direction, n)
    if direction == "c2c" and n > 4096 and four_step(n, batch):
        dev += 8 * n * batch
    return dev


def host_bytes_exact(direction, n, batch):
    """bfly_check's host buffers: input, output, the table, the R2C
    generator's complex scratch, and the per-row verify vectors."""
    host = input_bytes_for(direction, n, batch) + output_bytes_for(direction, n, batch) + table_bytes(direction, n)
    if direction == "r2c":
        host += 8 * n * batch
    return host + 32 * n


def budget_case(direction, n, batch, host_avail, device_free, device_total):
    dev = device_bytes_exact(direction, n, batch)
    host = host_bytes_exact(direction, n, batch)
    dev_headroom = max(4 * GIB, int(DEVICE_HEADROOM * device_total))
    host_headroom = max(2 * GIB, int(HOST_MARGIN * host_avail))
    out = {"input_bytes": input_bytes_for(direction, n, batch), "host_needed": host,
           "device_needed": dev, "decision": "run", "reason": ""}
    floats = max(input_bytes_for(direction, n, batch), output_bytes_for(direction, n, batch)) // 4
    if floats >= OFFSET_LIMIT:
        out.update(decision="skip", reason=f"{floats} float elements reach the 32-bit offset limit")
    elif dev > device_free - dev_headroom:
        out.update(decision="skip", reason=f"device needs {dev} B > free {device_free} B - headroom {dev_headroom} B")
    elif host > host_avail - host_headroom:
        out.update(decision="skip", reason=f"host needs {host} B > available {host_avail} B - headroom {host_headroom} B")
    return out


def plan(profile, host_avail, device_free, device_total):
    """Rows of the stress run: per tier, every C2C length (the canonical one
    as the soak), every R2C and C2R length."""
    c2c_ns = [int(n) for n in profile.get("c2c", {}).get("ns", [])]
    real = profile.get("real", {})
    tiers = [int(t) for t in profile.get("input_buffer_bytes", [])]
    soak_reps = int(profile.get("repeat_execution", 10000))
    soak_n = 1024 if 1024 in c2c_ns else (c2c_ns[len(c2c_ns) // 2] if c2c_ns else 0)
    rows = []

    def add(tier, stage, direction, n, reps, mode):
        batch = derive_batch(n, tier, direction)
        if batch:
            b = budget_case(direction, n, batch, host_avail, device_free, device_total)
        else:
            b = {"input_bytes": 0, "host_needed": "", "device_needed": "", "decision": "skip",
                 "reason": f"tier {tier} B cannot hold one n={n} {direction} input row"}
        row = dict.fromkeys(FIELDS, "")
        row.update(stage=stage, tier_bytes=tier, direction=direction, n=n, batch=batch, reps=reps,
                   input=mode, host_avail=host_avail, device_free=device_free,
                   device_total=device_total, **b)
        rows.append(row)

    for tier in tiers:
        for n in c2c_ns:
            soak = n == soak_n
            add(tier, "soak" if soak else "case", "c2c", n, soak_reps if soak else QUICK_REPS, "random-seeded")
        for n in real.get("r2c_ns", []):
            add(tier, "case", "r2c", int(n), QUICK_REPS, "random-seeded")
        for n in real.get("c2r_ns", []):
            add(tier, "case", "c2r", int(n), QUICK_REPS, "half-spectrum")
    return rows


def read_host_available(path="/proc/meminfo"):
    m = re.search(r"MemAvailable:\s+(\d+) kB", Path(path).read_text())
    return int(m.group(1)) * 1024 if m else 0


def read_device_memory(lib):
    import ctypes as C
    free, total = C.c_size_t(), C.c_size_t()
    rc = lib.bfly_rt_mem_info(C.byref(free), C.byref(total))
    if rc != 0:
        raise RuntimeError(f"bfly_rt_mem_info failed: {rc}")
    return free.value, total.value


def read_rss_bytes(pid):
    try:
        m = re.search(r"VmRSS:\s+(\d+) kB", Path(f"/proc/{pid}/status").read_text())
    except OSError:
        return None
    return int(m.group(1)) * 1024 if m else None


def rss_growth_bytes(samples):
    """Growth of the tail's peak over the first-quarter baseline, or None
    with fewer than RSS_MIN_SAMPLES samples."""
    if len(samples) < RSS_MIN_SAMPLES:
        return None
    base = samples[len(samples) // 4][1]
    tail = samples[-max(4, len(samples) // 10):]
    return max(0, max(r for _, r in tail) - base)


def rss_limit_bytes(peak):
    return max(96 * MIB, int(0.08 * peak))


def case_timeout_s(reps):
    return float(max(300, min(1800, 60 + reps // 10)))


def run_with_rss(argv, env, timeout_s):
    """Run argv sampling its RSS; the output goes to a temporary file (a pipe
    read only at the end could fill and stall the child)."""
    import tempfile
    t0 = time.perf_counter()
    with tempfile.TemporaryFile(mode="w+", errors="replace") as log:
        proc = subprocess.Popen(argv, env=env, stdout=log, stderr=subprocess.STDOUT)
        samples, timed_out = [], False
        while proc.poll() is None:
            if time.perf_counter() - t0 > timeout_s:
                timed_out = True
                proc.kill()
                break
            rss = read_rss_bytes(proc.pid)
            if rss is not None:
                samples.append((time.perf_counter() - t0, rss))
            time.sleep(RSS_SAMPLE_S)
        proc.wait()
        log.seek(0)
        out = log.read()
    return proc.returncode, out, samples, timed_out, (time.perf_counter() - t0) * 1e3


def _env(extra):
    env = {k: v for k, v in os.environ.items() if not k.startswith("AB_")}
    env.update(extra)
    return env


def execute_case(row, exe=DEFAULT_EXE):
    case = {"direction": row["direction"], "n": row["n"], "batch": row["batch"], "input": row["input"],
            "seed": 0}
    extra, argv = command_for(case, row["reps"], exe)
    timeout_s = case_timeout_s(row["reps"])
    rc, out, samples, timed_out, wall_ms = run_with_rss(argv, _env(extra), timeout_s)
    err, ok = verdict_of(rc, out)
    failure = f"timeout after {timeout_s:.0f} seconds" if timed_out else (
        "" if ok else "exit status, PASS line or maxRel check failed")
    ok = ok and not timed_out
    growth = ""
    if row["stage"] == "soak":
        measured = rss_growth_bytes(samples)
        if measured is not None:
            limit = rss_limit_bytes(max(r for _, r in samples))
            growth = round(measured / MIB, 3)
            if measured > limit:
                ok = False
                failure = f"rss grew {measured} B over the baseline, limit {limit} B, during the soak"
    row.update(returncode=rc, wall_ms=round(wall_ms, 1), max_rel="" if math.isnan(err) else err,
               rss_growth_mb=growth, correct=ok, failure=failure)
    return row


def run_lifecycle(row, cycles, exe=DEFAULT_EXE):
    """Start, run and end the program `cycles` times on one shape."""
    case = {"direction": row["direction"], "n": row["n"], "batch": row["batch"], "input": row["input"],
            "seed": 0}
    extra, argv = command_for(case, 1, exe)
    t0 = time.perf_counter()
    failed = []
    for cycle in range(1, cycles + 1):
        try:
            r = subprocess.run(argv, env=_env(extra), capture_output=True, text=True, timeout=120)
        except subprocess.TimeoutExpired:
            failed.append(cycle)
            continue
        if not verdict_of(r.returncode, r.stdout + r.stderr)[1]:
            failed.append(cycle)
    row.update(reps=cycles, returncode=0 if not failed else 1,
               wall_ms=round((time.perf_counter() - t0) * 1e3, 1), correct=not failed,
               failure="" if not failed else f"lifecycle cycles failed: {failed[:8]}")
    return row


def oom_probe(lib, device):
    """Refused oversize allocation, then a correct transform in the same
    process: returns (passed, detail).  The request is twice the device's
    memory, so the runtime refuses it without exhausting anything."""
    import ctypes as C

    import numpy as np
    if lib.bfly_rt_init(device) != 0:
        return False, "device init failed"
    free, total = read_device_memory(lib)
    p = C.c_void_p()
    rc = lib.bfly_rt_malloc(C.byref(p), 2 * total)
    if rc != -3 or p.value:
        return False, f"oversize allocation not refused as out of memory (rc {rc})"
    # recovery: allocate, transform an impulse row, check the flat spectrum
    n, kind = 64, 0
    table = np.empty(lib.bfly_table_floats(kind, n), np.float32)
    if lib.bfly_table_fill(kind, n, table.ctypes.data) != 0:
        return False, "table fill failed"
    x = np.zeros(2 * n, np.float32)
    x[0] = 1.0
    bufs = [C.c_void_p() for _ in range(3)]
    stream = C.c_void_p()
    try:
        if lib.bfly_rt_stream_create(C.byref(stream)) != 0:
            return False, "stream create failed after the refused allocation"
        for b, size in zip(bufs, (x.nbytes, x.nbytes, table.nbytes)):
            if lib.bfly_rt_malloc(C.byref(b), size) != 0:
                return False, "allocation failed after the refused one"
        y = np.empty_like(x)
        steps = (lib.bfly_rt_memcpy_h2d(bufs[2], table.ctypes.data, table.nbytes, stream),
                 lib.bfly_rt_memcpy_h2d(bufs[0], x.ctypes.data, x.nbytes, stream),
                 lib.bfly_c2c(bufs[0], bufs[1], n, 1, -1, bufs[2], None, stream),
                 lib.bfly_rt_memcpy_d2h(y.ctypes.data, bufs[1], y.nbytes, stream))
        if any(steps):
            return False, f"transform after the refused allocation failed {steps}"
        want = np.zeros(2 * n, np.float32)
        want[0::2] = 1.0
        if not np.allclose(y, want, atol=1e-6):
            return False, "wrong transform after the refused allocation"
        return True, f"refused {2 * total} B (free {free} B of {total} B), then recovered"
    finally:
        for b in bufs:
            lib.bfly_rt_free(b)
        lib.bfly_rt_stream_destroy(stream)


def print_plan(rows, soak_reps, host_avail, device_free, device_total):
    print(f"plan rows={len(rows)} soak_reps={soak_reps} host_avail={host_avail} "
          f"device_free={device_free} device_total={device_total}")
    for r in rows:
        print(f"  {r['stage']:9} tier={r['tier_bytes']:>10} {r['direction']} n={r['n']:<6} "
              f"batch={r['batch']:<9} reps={r['reps']:<6} {r['decision']} {r['reason']}")


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--config", default=str(DEFAULT_CONFIG))
    ap.add_argument("--exe", default=str(DEFAULT_EXE))
    ap.add_argument("--lib", choices=("hip", "stub"), default="hip",
                    help="library answering the device-memory query (stub: CPU runs)")
    ap.add_argument("--plan", action="store_true", help="print the budgeted plan and exit")
    ap.add_argument("--soak-reps", type=int, default=None)
    ap.add_argument("--lifecycle-cycles", type=int, default=20)
    ap.add_argument("--tiers", default="", help="comma-separated subset of the profile's tiers (bytes)")
    ap.add_argument("--oom-probe", action="store_true",
                    help="opt-in: refused oversize allocation, then recovery in the same process")
    ap.add_argument("--device", type=int, default=None, help="device of the OOM probe (required with it)")
    ap.add_argument("--out", default="")
    args = ap.parse_args(argv)
    if args.oom_probe and args.device is None:
        ap.error("--oom-probe requires an explicit --device")
    if args.soak_reps is not None and args.soak_reps < 1:
        ap.error("--soak-reps must be positive")
    if args.lifecycle_cycles < 1:
        ap.error("--lifecycle-cycles must be positive")
    profile = dict(load_config(args.config)["profiles"]["stress"])
    if args.soak_reps is not None:
        profile["repeat_execution"] = args.soak_reps
    if args.tiers:
        keep = {int(t) for t in args.tiers.split(",")}
        profile["input_buffer_bytes"] = [t for t in profile["input_buffer_bytes"] if int(t) in keep]
    from bfly import _lib
    lib = _lib.hip() if args.lib == "hip" else _lib.stub()
    if args.lib == "hip" and lib.bfly_rt_init(args.device or 0) != 0:
        ap.error("no device")
    device_free, device_total = read_device_memory(lib)
    host_avail = read_host_available()
    rows = plan(profile, host_avail, device_free, device_total)
    soak_reps = int(profile.get("repeat_execution", 10000))
    print_plan(rows, soak_reps, host_avail, device_free, device_total)
    if args.plan:
        return 0
    if not Path(args.exe).is_file():
        ap.error(f"{args.exe} is missing (make -C ascend-fft_amd)")
    stamp = datetime.now(timezone.utc).strftime("%Y%m%dT%H%M%SZ")
    out = Path(args.out) if args.out else Path.cwd() / "bfly_runs" / f"{stamp}-stress"
    out.mkdir(parents=True, exist_ok=True)
    done, lifecycle_from = [], None
    runnable = [r for r in rows if r["decision"] == "run"]
    for row in rows:
        if row["decision"] != "run":
            done.append(row)
            continue
        print(f"[{sum(1 for d in done if d['decision'] == 'run') + 1}/{len(runnable)}] {row['stage']} "
              f"{row['direction']} n={row['n']} batch={row['batch']} reps={row['reps']}", flush=True)
        done.append(execute_case(row, args.exe))
        if not row["correct"]:
            print(f"  FAILED: {row['failure']}", flush=True)
        if lifecycle_from is None and row["stage"] == "case" and row["direction"] == "c2c" and row["correct"]:
            lifecycle_from = row
    if lifecycle_from is not None:
        lc = dict.fromkeys(FIELDS, "")
        lc.update({k: lifecycle_from[k] for k in ("tier_bytes", "direction", "n", "batch", "input")})
        lc.update(stage="lifecycle", decision="run")
        print(f"lifecycle: {args.lifecycle_cycles} cycles of n={lc['n']} batch={lc['batch']}", flush=True)
        done.append(run_lifecycle(lc, args.lifecycle_cycles, args.exe))
    if args.oom_probe:
        ok, detail = oom_probe(lib, args.device)
        probe = dict.fromkeys(FIELDS, "")
        probe.update(stage="oom_probe", decision="run", device_total=device_total, returncode=0 if ok else 1,
                     correct=ok, failure="" if ok else detail, reason=detail)
        print(f"oom probe: {'PASS' if ok else 'FAIL'} ({detail})", flush=True)
        done.append(probe)
    with (out / "stress.csv").open("w", newline="", encoding="utf-8") as fh:
        w = csv.DictWriter(fh, fieldnames=FIELDS)
        w.writeheader()
        w.writerows(done)
    ran = [r for r in done if r["decision"] == "run"]
    failed = [r for r in ran if not r["correct"]]
    soaks = [r for r in ran if r["stage"] == "soak"]
    summary = {"schema_version": 1, "profile": "stress", "rows": len(done), "executed": len(ran),
               "skipped": len(done) - len(ran), "passed": len(ran) - len(failed), "failed": len(failed),
               "soak_reps": soak_reps, "soak_rss_growth_mb": [r["rss_growth_mb"] for r in soaks],
               "lifecycle_cycles": args.lifecycle_cycles if lifecycle_from is not None else 0,
               "oom_probe": args.oom_probe, "status": "pass" if not failed else "fail"}
    (out / "summary.json").write_text(json.dumps(summary, indent=2) + "\n", encoding="utf-8")
    print(f"{summary['passed']}/{summary['executed']} PASS, {summary['skipped']} skipped by budget -> {out}")
    return 0 if not failed else 1


if __name__ == "__main__":
    sys.exit(main())
