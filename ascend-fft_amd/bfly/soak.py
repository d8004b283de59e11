"""Memory-budgeted soak, lifecycle and allocation-failure checks: the runner
the `stress` profile of config/test_matrix.json names.

It covers what the reference's scripts/run_stress.py covers (declared input
tiers, a budget before any allocation, a long repeated-execution soak with a
resident-memory watch, process lifecycles, an opt-in allocation-failure
probe), on this library's own plan:

* Each tier T of `input_buffer_bytes` gives, per declared length, the batch
  whose input fills T.  A job is admitted only when its exact footprint fits:
  the device bytes bfly_check allocates (input, output, the (kind, n) table
  from bfly_table_floats, the four-step workspace when bfly_c2c_launches says
  the chain runs) under the free device memory less max(4 GiB, 5 %), the host
  bytes under MemAvailable less max(2 GiB, 2 %), and fewer than 2^32 floats
  per buffer.  A refused job keeps its figures and the reason.
* Per tier, the soak length (1024 when declared) runs `repeat_execution`
  executions in one process while its RSS is sampled; growth of the late
  samples over the early ones beyond max(96 MiB, 8 % of the peak) fails it.
* The lifecycle stage starts, runs and ends bfly_check `--cycles` times on
  the first C2C job that passed.
* `--oom-probe` (with an explicit `--device`) asks for twice the device's
  memory through bfly_rt_malloc, which the runtime refuses without
  exhausting anything, then allocates and transforms an impulse correctly in
  the same process.

Executions are classified as in bfly.suite.  A wrong answer is recorded and
the run goes on; the first abnormal end (signal, time limit, no verdict line)
stops everything: no further job, lifecycle cycle or probe runs.

    python -m bfly.soak [--plan] [--soak-reps K] [--cycles C] [--tiers B,..]
                        [--oom-probe --device D] [--lib hip|stub] [--out DIR]
"""
from __future__ import annotations

import argparse
import csv
import ctypes as C
import json
import math
import re
import sys
from dataclasses import dataclass
from pathlib import Path
from typing import List, Optional, Tuple

PKG = Path(__file__).resolve().parents[1]
if str(PKG) not in sys.path:
    sys.path.insert(0, str(PKG))

from bfly import _lib  # noqa: E402
from bfly.suite import CHECK_EXE, MATRIX, Case, Outcome, default_out, execute, load_profiles  # noqa: E402

MIB, GIB = 1 << 20, 1 << 30
FLOAT_OFFSET_LIMIT = 1 << 32
CASE_REPS = 3
KINDS = {"c2c": _lib.KIND_C2C, "r2c": _lib.KIND_R2C, "c2r": _lib.KIND_C2R}


# ------------------------------------------------------------------ memory
def row_floats(kind: str, n: int) -> Tuple[int, int]:
    """(input, output) floats of one row: interleaved complex, n reals, n/2+1 bins."""
    return {"c2c": (2 * n, 2 * n), "r2c": (n, n + 2), "c2r": (n + 2, n)}[kind]


def batch_for_tier(kind: str, n: int, tier_bytes: int) -> int:
    return tier_bytes // (4 * row_floats(kind, n)[0]) if n > 0 and tier_bytes > 0 else 0


def _shape_lib():
    """Answers table and launch-count queries without touching a device: the
    gfx950 library when it is built, else its CPU stub (same host code)."""
    try:
        return _lib.hip()
    except RuntimeError:
        return _lib.stub()


@dataclass(frozen=True)
class Footprint:
    kind: str
    n: int
    batch: int
    device: int
    host: int
    floats: int          # the larger buffer, in floats

    @classmethod
    def of(cls, kind: str, n: int, batch: int) -> "Footprint":
        lib = _shape_lib()
        fin, fout = (f * batch for f in row_floats(kind, n))
        table = 4 * int(lib.bfly_table_floats(KINDS[kind], n))
        work = 8 * n * batch if kind == "c2c" and lib.bfly_c2c_launches(n, batch, 0) == 2 else 0
        device = 4 * (fin + fout) + table + work
        # bfly_check's host buffers: input, output, table, per-row check
        # vectors, and the R2C generator's complex scratch
        host = 4 * (fin + fout) + table + 32 * n + (8 * n * batch if kind == "r2c" else 0)
        return cls(kind, n, batch, device, host, max(fin, fout))


@dataclass(frozen=True)
class Budget:
    host_available: int
    device_free: int
    device_total: int

    @property
    def device_reserve(self) -> int:
        return max(4 * GIB, self.device_total // 20)

    @property
    def host_reserve(self) -> int:
        return max(2 * GIB, self.host_available // 50)

    def refusal(self, fp: Footprint) -> str:
        """Why the footprint does not fit ('' when it does)."""
        if fp.floats >= FLOAT_OFFSET_LIMIT:
            return f"a buffer of {fp.floats} floats needs offsets beyond 32 bits"
        if fp.device > self.device_free - self.device_reserve:
            return (f"device: {fp.device} B wanted, {self.device_free} B free less a "
                    f"{self.device_reserve} B reserve")
        if fp.host > self.host_available - self.host_reserve:
            return (f"host: {fp.host} B wanted, {self.host_available} B available less a "
                    f"{self.host_reserve} B reserve")
        return ""


# -------------------------------------------------------------------- plan
@dataclass
class Job:
    stage: str            # case | soak | lifecycle | oom_probe
    tier: int
    case: Optional[Case]
    reps: int
    footprint: Optional[Footprint]
    refusal: str
    outcome: Optional[Outcome] = None
    rss_growth: Optional[int] = None

    @property
    def admitted(self) -> bool:
        return not self.refusal


def plan_jobs(spec: dict, budget: Budget, soak_reps: Optional[int] = None) -> List[Job]:
    """Per tier: every C2C length (the soak length as the soak), then every
    R2C and C2R length, each budgeted."""
    lengths = {k: [int(n) for n in v] for k, v in (("c2c", spec.get("c2c", {}).get("ns", [])),
                                                     ("r2c", spec.get("real", {}).get("r2c_ns", [])),
                                                     ("c2r", spec.get("real", {}).get("c2r_ns", [])))}
    soak_n = 1024 if 1024 in lengths["c2c"] else (sorted(lengths["c2c"])[len(lengths["c2c"]) // 2]
                                                   if lengths["c2c"] else None)
    soak_reps = int(spec.get("repeat_execution", 10000) if soak_reps is None else soak_reps)
    jobs = []
    for tier in (int(t) for t in spec.get("input_buffer_bytes", [])):
        for kind in ("c2c", "r2c", "c2r"):
            for n in lengths[kind]:
                soak = kind == "c2c" and n == soak_n
                batch = batch_for_tier(kind, n, tier)
                case = Case("soak", kind, n, batch, "half-spectrum" if kind == "c2r" else "random-seeded")
                if batch == 0:
                    jobs.append(Job("soak" if soak else "case", tier, case, 0, None,
                                    f"a {tier} B tier holds no {kind} row of {n} points"))
                    continue
                fp = Footprint.of(kind, n, batch)
                jobs.append(Job("soak" if soak else "case", tier, case, soak_reps if soak else CASE_REPS,
                                fp, budget.refusal(fp)))
    return jobs


# ---------------------------------------------------------------- rss watch
def rss_growth(samples: List[Tuple[float, int]]) -> Optional[int]:
    """Peak of the last tenth of the samples (at least 4) over the sample a
    quarter of the way in; None with fewer than 8 samples."""
    if len(samples) < 8:
        return None
    early = samples[len(samples) // 4][1]
    late = max(r for _, r in samples[-max(4, len(samples) // 10):])
    return max(0, late - early)


def rss_allowance(peak: int) -> int:
    return max(96 * MIB, (8 * peak) // 100)


def time_limit_s(reps: int) -> float:
    return float(min(1800, max(300, 60 + reps // 10)))


# ---------------------------------------------------------------- the probe
def oom_probe(lib, device: int) -> Tuple[bool, str]:
    """An oversize request refused as out of memory, then a correct 64-point
    impulse transform in the same process."""
    if lib.bfly_rt_init(device) != 0:
        return False, "device init failed"
    free, total = device_memory(lib)
    p = C.c_void_p()
    rc = lib.bfly_rt_malloc(C.byref(p), 2 * total)
    if rc != _lib.ERR_ALLOC or p.value:
        return False, f"a request for twice the device memory returned {rc}, not an allocation refusal"
    n = 64
    import numpy as np
    table = np.empty(lib.bfly_table_floats(_lib.KIND_C2C, n), np.float32)
    x = np.zeros(2 * n, np.float32)
    x[0] = 1.0                                # impulse: every bin is 1 + 0i
    y = np.full_like(x, np.nan)
    ptrs = [C.c_void_p() for _ in range(3)]
    stream = C.c_void_p()
    try:
        steps = [("table", lib.bfly_table_fill(_lib.KIND_C2C, n, table.ctypes.data)),
                 ("stream", lib.bfly_rt_stream_create(C.byref(stream)))]
        steps += [("malloc", lib.bfly_rt_malloc(C.byref(q), size))
                  for q, size in zip(ptrs, (x.nbytes, x.nbytes, table.nbytes))]
        if any(rc for _, rc in steps):
            return False, f"setup after the refusal failed: {steps}"
        steps = [("h2d table", lib.bfly_rt_memcpy_h2d(ptrs[2], table.ctypes.data, table.nbytes, stream)),
                 ("h2d input", lib.bfly_rt_memcpy_h2d(ptrs[0], x.ctypes.data, x.nbytes, stream)),
                 ("c2c", lib.bfly_c2c(ptrs[0], ptrs[1], n, 1, -1, ptrs[2], None, stream)),
                 ("d2h", lib.bfly_rt_memcpy_d2h(y.ctypes.data, ptrs[1], y.nbytes, stream))]
        if any(rc for _, rc in steps):
            return False, f"transform after the refusal failed: {steps}"
        if not (np.abs(y[0::2] - 1.0).max() <= 1e-6 and np.abs(y[1::2]).max() <= 1e-6):
            return False, "the transform after the refusal gave a wrong spectrum"
        return True, f"{2 * total} B refused ({free} B free of {total} B), then an impulse transformed correctly"
    finally:
        for q in ptrs:
            if q.value:
                lib.bfly_rt_free(q)
        if stream.value:
            lib.bfly_rt_stream_destroy(stream)


def device_memory(lib) -> Tuple[int, int]:
    free, total = C.c_size_t(), C.c_size_t()
    rc = lib.bfly_rt_mem_info(C.byref(free), C.byref(total))
    if rc != 0:
        raise RuntimeError(f"bfly_rt_mem_info returned {rc}")
    return free.value, total.value


def host_available(meminfo="/proc/meminfo") -> int:
    m = re.search(r"MemAvailable:\s+(\d+) kB", Path(meminfo).read_text())
    return int(m.group(1)) * 1024 if m else 0


# -------------------------------------------------------------------- run
ROW_FIELDS = ("stage", "tier", "kind", "n", "batch", "reps", "input", "device_bytes", "host_bytes",
              "admitted", "refusal", "status", "returncode", "max_rel", "wall_ms", "rss_growth_mib", "note")


def _row(job: Job) -> dict:
    c, fp, o = job.case, job.footprint, job.outcome
    return {"stage": job.stage, "tier": job.tier or "", "kind": c.kind if c else "", "n": c.n if c else "",
            "batch": c.batch if c else "", "reps": job.reps, "input": c.input if c else "",
            "device_bytes": fp.device if fp else "", "host_bytes": fp.host if fp else "",
            "admitted": job.admitted, "refusal": job.refusal,
            "status": o.status if o else ("refused" if not job.admitted else ""),
            "returncode": o.returncode if o and o.returncode is not None else "",
            "max_rel": o.max_rel if o and math.isfinite(o.max_rel) else "",
            "wall_ms": round(o.wall_ms, 1) if o and math.isfinite(o.wall_ms) else "",
            "rss_growth_mib": round(job.rss_growth / MIB, 3) if job.rss_growth is not None else "",
            "note": o.note if o else ""}


def run(jobs: List[Job], exe, cycles: int, probe_device: Optional[int], lib, say=print) -> List[Job]:
    """Execute the admitted jobs, then the lifecycle and the probe; an
    abnormal end stops every later step (recorded not-run)."""
    halt = ""
    admitted = [j for j in jobs if j.admitted]
    for i, job in enumerate(admitted, 1):
        if halt:
            job.outcome = Outcome("not-run", note=halt)
            continue
        say(f"[{i}/{len(admitted)}] {job.stage} tier={job.tier} {job.case.label} reps={job.reps}")
        o = execute(job.case, job.reps, exe, limit_s=time_limit_s(job.reps), sample_rss=job.stage == "soak")
        job.outcome = o
        if job.stage == "soak" and o.status == "pass":
            job.rss_growth = rss_growth(o.rss)
            if job.rss_growth is not None:
                allowed = rss_allowance(max(r for _, r in o.rss))
                if job.rss_growth > allowed:
                    o.status, o.note = "wrong", (f"resident memory grew {job.rss_growth} B during the soak "
                                                 f"(allowed {allowed} B)")
        if o.status == "abnormal":
            halt = f"not run: {job.case.label} ended abnormally ({o.note})"
            say(f"  ABNORMAL: {o.note}; nothing further runs")
        elif o.status == "wrong":
            say(f"  WRONG: {o.note}")
    extra = []
    first_c2c = next((j for j in admitted if j.stage == "case" and j.case.kind == "c2c"
                      and j.outcome and j.outcome.passed), None)
    if first_c2c is not None and cycles > 0:
        life = Job("lifecycle", first_c2c.tier, first_c2c.case, cycles, first_c2c.footprint, "")
        if halt:
            life.outcome = Outcome("not-run", note=halt)
        else:
            say(f"lifecycle: {cycles} start-run-end cycles of {first_c2c.case.label}")
            wall, failed = 0.0, []
            for k in range(1, cycles + 1):
                o = execute(first_c2c.case, 1, exe, limit_s=120.0)
                wall += o.wall_ms
                if o.status == "abnormal":
                    halt = f"not run: lifecycle cycle {k} ended abnormally ({o.note})"
                    life.outcome = Outcome("abnormal", o.returncode, note=f"cycle {k}: {o.note}", wall_ms=wall)
                    break
                if not o.passed:
                    failed.append(k)
            if life.outcome is None:
                life.outcome = Outcome("pass" if not failed else "wrong", 0 if not failed else 1, wall_ms=wall,
                                       note="" if not failed else f"cycles with a wrong answer: {failed[:8]}")
        extra.append(life)
    if probe_device is not None:
        probe = Job("oom_probe", 0, None, 1, None, "")
        if halt:
            probe.outcome = Outcome("not-run", note=halt)
        else:
            ok, detail = oom_probe(lib, probe_device)
            probe.outcome = Outcome("pass" if ok else "wrong", 0 if ok else 1, note=detail)
            say(f"oom probe: {'pass' if ok else 'FAILED'} ({detail})")
        extra.append(probe)
    return jobs + extra


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(prog="python -m bfly.soak", description=__doc__,
                                 formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--matrix", default=str(MATRIX))
    ap.add_argument("--exe", default=str(CHECK_EXE))
    ap.add_argument("--lib", choices=("hip", "stub"), default="hip",
                    help="library answering the device-memory query (stub: CPU-only runs)")
    ap.add_argument("--plan", action="store_true", help="print the budgeted plan and exit")
    ap.add_argument("--soak-reps", type=int, default=None)
    ap.add_argument("--cycles", type=int, default=20, help="lifecycle cycles")
    ap.add_argument("--tiers", default="", help="comma-separated subset of the declared tiers (bytes)")
    ap.add_argument("--oom-probe", action="store_true", help="opt in to the allocation-failure probe")
    ap.add_argument("--device", type=int, default=None, help="device of the probe (required with it)")
    ap.add_argument("--out", default="")
    args = ap.parse_args(argv)
    if args.oom_probe and args.device is None:
        ap.error("--oom-probe needs an explicit --device")
    if args.soak_reps is not None and args.soak_reps < 1:
        ap.error("--soak-reps takes a positive count")
    if args.cycles < 0:
        ap.error("--cycles takes a count >= 0")
    spec = dict(load_profiles(args.matrix)["stress"].spec)
    if args.tiers:
        keep = {int(t) for t in args.tiers.split(",")}
        spec["input_buffer_bytes"] = [t for t in spec.get("input_buffer_bytes", []) if int(t) in keep]
    lib = _lib.hip() if args.lib == "hip" else _lib.stub()
    if args.lib == "hip" and lib.bfly_rt_init(args.device or 0) != 0:
        ap.error("no device")
    free, total = device_memory(lib)
    budget = Budget(host_available(), free, total)
    jobs = plan_jobs(spec, budget, args.soak_reps)
    print(f"plan: {len(jobs)} jobs, {sum(j.admitted for j in jobs)} admitted; host available "
          f"{budget.host_available} B, device free {free} B of {total} B")
    for j in jobs:
        print(f"  {j.stage:5} tier={j.tier:<10} {j.case.kind} n={j.case.n:<6} batch={j.case.batch:<9} "
              f"reps={j.reps:<6} {'admitted' if j.admitted else 'refused: ' + j.refusal}")
    if args.plan:
        return 0
    if not Path(args.exe).is_file():
        ap.error(f"{args.exe} does not exist (build it: make -C ascend-fft_amd)")
    done = run(jobs, args.exe, args.cycles, args.device if args.oom_probe else None, lib,
               say=lambda s: print(s, flush=True))
    out = Path(args.out) if args.out else default_out("stress")
    out.mkdir(parents=True, exist_ok=True)
    rows = [_row(j) for j in done]
    with (out / "stress.csv").open("w", newline="", encoding="utf-8") as fh:
        w = csv.DictWriter(fh, fieldnames=ROW_FIELDS)
        w.writeheader()
        w.writerows(rows)
    ran = [r for r in rows if r["status"] not in ("refused", "")]
    count = {s: sum(r["status"] == s for r in ran) for s in ("pass", "wrong", "abnormal", "not-run")}
    life = next((j for j in done if j.stage == "lifecycle"), None)
    summary = {"profile": "stress", "jobs": len(rows), "refused": sum(r["status"] == "refused" for r in rows),
               **{k.replace("-", "_"): v for k, v in count.items()},
               "soak_reps": next((j.reps for j in done if j.stage == "soak"), 0),
               "soak_rss_growth_mib": [r["rss_growth_mib"] for r in rows if r["stage"] == "soak" and r["status"]
                                       not in ("refused", "not-run")],
               "lifecycle_cycles": life.reps if life and life.outcome and life.outcome.passed else 0,
               "oom_probe": bool(args.oom_probe),
               "result": "pass" if ran and count["pass"] == len(ran) else "fail"}
    (out / "summary.json").write_text(json.dumps(summary, indent=2) + "\n", encoding="utf-8")
    print(f"{count['pass']}/{len(ran)} pass, {count['wrong']} wrong, {count['abnormal']} abnormal, "
          f"{count['not-run']} not run, {summary['refused']} refused by the budget -> {out}")
    return 0 if summary["result"] == "pass" else 1


if __name__ == "__main__":
    sys.exit(main())
