"""Correctness suites of config/test_matrix.json, run case by case through
the `bfly_check` program.

The profile file keeps the reference's schema (config/test_matrix.json of
the reference: profile names, `c2c` grid and `fixed_total_points`, `real`
lengths, `numeric_patterns` / `numeric_cases` / `random_seeds`,
`applications`, `runnable`, `runner`), so a profile written for the
reference's runner (scripts/run_test_profile.py there) runs here.  The
runner itself is this library's:

* a profile expands into `Case` records through a table of components
  (c2c grid, real rows, numerical patterns, application shapes);
* every execution ends in one of four outcomes: ``pass``; ``wrong`` (the
  program exited normally but its verdict failed the gate: exit status 0, a
  PASS line and 0 <= maxRel <= 1e-4 are all required); ``abnormal`` (killed
  by a signal, timed out, or exited without any verdict line: a fault or a
  hang, not an answer); ``not-run``;
* a wrong answer is recorded and the run goes on; the first abnormal outcome
  ends the run, and every case after it is recorded ``not-run`` -- a
  faulting or hanging kernel is never launched again by the runner.

Outputs: ``cases.csv`` (one row per execution) and ``summary.json``.  The
`stress` profile names its own runner (`bfly.soak`), which shares the
execution and classification code below.

    python -m bfly.suite <profile> [--components c2c,real,numeric,applications]
                         [--reps K] [--trials T] [--dry-run] [--list]
                         [--out DIR] [--exe PATH]
"""
from __future__ import annotations

import argparse
import csv
import json
import math
import os
import re
import subprocess
import sys
import tempfile
import time
from dataclasses import asdict, dataclass, field, replace
from datetime import datetime, timezone
from pathlib import Path
from typing import Callable, Dict, Iterable, Iterator, List, Optional, Sequence, Tuple

PKG = Path(__file__).resolve().parents[1]
MATRIX = PKG / "config" / "test_matrix.json"
CHECK_EXE = PKG / "lib" / "bfly_check"
GATE = 1e-4
# application ids of the profile schema -> bfly_check input modes
APPLICATION_INPUTS = {"ofdm": "ofdm", "radar-range": "radar", "dl-frequency-layer": "dl", "stft": "sin"}
# default trials: profiles whose worst error over repeated runs is recorded
REPEATED_PROFILES = {"regression": 3, "publication": 3}


# ----------------------------------------------------------------- profiles
@dataclass(frozen=True)
class Profile:
    name: str
    spec: dict

    @property
    def runnable(self) -> bool:
        return bool(self.spec.get("runnable"))

    @property
    def runner(self) -> Optional[str]:
        return self.spec.get("runner") or None

    @property
    def description(self) -> str:
        return self.spec.get("description", "")


def load_profiles(path=MATRIX) -> Dict[str, Profile]:
    """Profiles of a test-matrix file; ValueError unless it is schema 1."""
    doc = json.loads(Path(path).read_text(encoding="utf-8"))
    if not isinstance(doc, dict) or doc.get("schema_version") != 1 or not isinstance(doc.get("profiles"), dict):
        raise ValueError(f"{path}: not a schema-1 test matrix")
    return {name: Profile(name, spec) for name, spec in doc["profiles"].items()}


@dataclass(frozen=True, order=True)
class Case:
    """One bfly_check invocation: transform kind, shape, input mode and seed."""
    group: str
    kind: str
    n: int
    batch: int
    input: str
    seed: int = 0

    def env(self) -> Dict[str, str]:
        env = {"AB_DIR": self.kind}
        if self.input != "half-spectrum":       # bfly_check's own C2R input
            env["AB_INPUT"] = self.input
        if self.seed:
            env["AB_SEED"] = str(self.seed)
        return env

    def argv(self, exe, reps: int) -> List[str]:
        return [str(exe), str(self.n), str(self.batch), str(reps)]

    def shell(self, exe, reps: int) -> str:
        return " ".join(f"{k}={v}" for k, v in sorted(self.env().items())) + " " + " ".join(self.argv(exe, reps))

    @property
    def label(self) -> str:
        seed = f" seed={self.seed}" if self.seed else ""
        return f"{self.group} {self.kind} n={self.n} batch={self.batch} input={self.input}{seed}"


def c2c_shapes(spec: dict) -> List[Tuple[int, int]]:
    """(n, batch) of the C2C grid and of the equal-total-point shapes."""
    grid = spec.get("c2c", {})
    ns = [int(n) for n in grid.get("ns", [])]
    shapes = {(n, int(b)) for n in ns for b in grid.get("batches", [])}
    for total in grid.get("fixed_total_points", []):
        shapes |= {(n, int(total) // n) for n in ns if int(total) % n == 0}
    return sorted(shapes)


def _c2c(spec: dict) -> Iterator[Case]:
    for n, b in c2c_shapes(spec):
        yield Case("c2c", "c2c", n, b, "sin")


def _real(spec: dict) -> Iterator[Case]:
    real = spec.get("real", {})
    for kind, lengths, source in (("r2c", "r2c_ns", "sin"), ("c2r", "c2r_ns", "half-spectrum")):
        for n in real.get(lengths, []):
            for b in real.get("batches", []):
                yield Case("real", kind, int(n), int(b), source)


def _numeric(spec: dict) -> Iterator[Case]:
    seeds = [int(s) for s in spec.get("random_seeds", [0])]
    for n, b in spec.get("numeric_cases", []):
        for pattern in spec.get("numeric_patterns", []):
            for seed in (seeds if pattern == "random-seeded" else [0]):
                yield Case("numeric", "c2c", int(n), int(b), pattern, seed)


def _applications(spec: dict) -> Iterator[Case]:
    for app in spec.get("applications", []):
        mode = APPLICATION_INPUTS.get(app.get("id"))
        if mode is None or app.get("transform") not in ("c2c", "r2c"):
            continue                              # 2-D shapes have no bfly_check mode
        for n in app.get("ns", []):
            for b in app.get("batches", []):
                yield Case("application", app["transform"], int(n), int(b), mode)


COMPONENTS: Dict[str, Callable[[dict], Iterator[Case]]] = {
    "c2c": _c2c, "real": _real, "numeric": _numeric, "applications": _applications,
}


def cases_of(profile: Profile, components: Iterable[str] = tuple(COMPONENTS)) -> List[Case]:
    """The profile's cases for the chosen components, sorted, each once."""
    unknown = sorted(set(components) - set(COMPONENTS))
    if unknown:
        raise ValueError(f"unknown components: {', '.join(unknown)}")
    return sorted({c for comp in components for c in COMPONENTS[comp](profile.spec)})


# ---------------------------------------------------------------- execution
@dataclass
class Outcome:
    status: str                      # pass | wrong | abnormal | not-run
    returncode: Optional[int] = None
    max_rel: float = math.nan
    wall_ms: float = math.nan
    device_us: float = math.nan
    note: str = ""
    tail: str = ""
    rss: List[Tuple[float, int]] = field(default_factory=list, repr=False)

    @property
    def passed(self) -> bool:
        return self.status == "pass"


def classify(returncode: Optional[int], output: str, timed_out: bool = False) -> Outcome:
    """Outcome of one finished bfly_check run (wall and RSS filled by the caller)."""
    verdicts = re.findall(r"^(PASS|FAIL)$", output, re.M)
    err = re.search(r"maxRel=(\S+)", output)
    try:
        max_rel = float(err.group(1)) if err else math.nan
    except ValueError:
        max_rel = math.nan
    chain = re.search(r"device_chain=([0-9.]+) us", output)
    o = Outcome("wrong", returncode, max_rel, device_us=float(chain.group(1)) if chain else math.nan)
    if timed_out:
        o.status, o.note = "abnormal", "no exit before the time limit"
    elif returncode is None or returncode < 0:
        o.status, o.note = "abnormal", f"ended by signal {-(returncode or 0)}"
    elif not verdicts:
        o.status, o.note = "abnormal", f"exit status {returncode} without a verdict line"
    elif returncode == 0 and verdicts[-1] == "PASS" and math.isfinite(max_rel) and 0.0 <= max_rel <= GATE:
        o.status = "pass"
    else:
        o.note = (f"verdict {verdicts[-1]}, exit status {returncode}, maxRel {max_rel:g} "
                  f"(needs PASS, 0 and a finite maxRel <= {GATE:g})")
    if o.status != "pass":
        o.tail = "\n".join(output.splitlines()[-12:])
    return o


def _rss(pid: int) -> Optional[int]:
    try:
        m = re.search(r"VmRSS:\s+(\d+) kB", Path(f"/proc/{pid}/status").read_text())
    except OSError:
        return None
    return int(m.group(1)) * 1024 if m else None


def execute(case: Case, reps: int, exe=CHECK_EXE, limit_s: float = 600.0, sample_rss: bool = False,
            every_s: float = 0.02) -> Outcome:
    """Run one case as a child process (output through a temporary file, so a
    full pipe never stalls it), optionally sampling its RSS, and classify."""
    env = {k: v for k, v in os.environ.items() if not k.startswith("AB_")}
    env.update(case.env())
    t0 = time.perf_counter()
    samples: List[Tuple[float, int]] = []
    timed_out = False
    with tempfile.TemporaryFile(mode="w+", errors="replace") as log:
        child = subprocess.Popen(case.argv(exe, reps), env=env, stdout=log, stderr=subprocess.STDOUT)
        while True:
            try:
                child.wait(timeout=every_s if sample_rss else 0.5)
                break
            except subprocess.TimeoutExpired:
                pass
            if sample_rss:
                r = _rss(child.pid)
                if r is not None:
                    samples.append((time.perf_counter() - t0, r))
            if time.perf_counter() - t0 > limit_s:
                timed_out = True
                child.kill()
                child.wait()
                break
        log.seek(0)
        text = log.read()
    o = classify(child.returncode, text, timed_out)
    o.wall_ms = (time.perf_counter() - t0) * 1e3
    o.rss = samples
    return o


ROW_FIELDS = ("group", "kind", "n", "batch", "input", "seed", "trial", "status", "returncode", "max_rel",
              "wall_ms", "device_us", "note", "tail")


def run_cases(cases: Sequence[Case], trials: int, reps: int, exe=CHECK_EXE,
              say: Callable[[str], None] = print) -> List[dict]:
    """Every (trial, case) in order; the first abnormal outcome ends the run
    and the remaining executions are recorded as not-run."""
    rows, stopped = [], ""
    total = len(cases) * trials
    for trial in range(1, trials + 1):
        for case in cases:
            if stopped:
                o = Outcome("not-run", note=stopped)
            else:
                say(f"[{len(rows) + 1}/{total}] trial {trial}: {case.label}")
                o = execute(case, reps, exe)
                if o.status == "abnormal":
                    stopped = f"not run: an earlier execution ended abnormally ({case.label}: {o.note})"
                    say(f"  ABNORMAL: {o.note}; stopping, the remaining executions are not run")
                elif o.status == "wrong":
                    say(f"  WRONG: {o.note}")
            row = dict(asdict(case), trial=trial, status=o.status, returncode=o.returncode, max_rel=o.max_rel,
                       wall_ms=round(o.wall_ms, 2) if math.isfinite(o.wall_ms) else "",
                       device_us=o.device_us, note=o.note, tail=o.tail)
            rows.append(row)
    return rows


def write_outputs(out: Path, profile: str, components, cases, trials: int, rows: List[dict]) -> dict:
    out.mkdir(parents=True, exist_ok=True)
    with (out / "cases.csv").open("w", newline="", encoding="utf-8") as fh:
        w = csv.DictWriter(fh, fieldnames=ROW_FIELDS)
        w.writeheader()
        w.writerows(rows)
    count = {s: sum(r["status"] == s for r in rows) for s in ("pass", "wrong", "abnormal", "not-run")}
    errors = [r["max_rel"] for r in rows if isinstance(r["max_rel"], float) and math.isfinite(r["max_rel"])]
    summary = {"profile": profile, "components": list(components), "cases": len(cases), "trials": trials,
               "executions": len(rows), **{k.replace("-", "_"): v for k, v in count.items()},
               "gate_max_rel": GATE, "worst_max_rel": max(errors) if errors else None,
               "result": "pass" if count["pass"] == len(rows) else "fail"}
    (out / "summary.json").write_text(json.dumps(summary, indent=2) + "\n", encoding="utf-8")
    return summary


def default_out(profile: str) -> Path:
    return Path.cwd() / "bfly_runs" / f"{datetime.now(timezone.utc):%Y%m%dT%H%M%SZ}-{profile}"


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(prog="python -m bfly.suite", description=__doc__,
                                 formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("profile", nargs="?", default="smoke")
    ap.add_argument("--matrix", default=str(MATRIX), help="test-matrix file")
    ap.add_argument("--exe", default=str(CHECK_EXE))
    ap.add_argument("--list", action="store_true", help="list the profiles and exit")
    ap.add_argument("--components", default=",".join(COMPONENTS))
    ap.add_argument("--trials", type=int, default=None)
    ap.add_argument("--reps", type=int, default=1)
    ap.add_argument("--dry-run", action="store_true", help="print the commands and exit")
    ap.add_argument("--out", default="")
    raw = list(sys.argv[1:] if argv is None else argv)
    args = ap.parse_args(raw)
    profiles = load_profiles(args.matrix)
    if args.list:
        for p in profiles.values():
            kind = f"runner {p.runner}" if p.runner else ("runnable" if p.runnable else "not runnable")
            print(f"{p.name:12} {kind:18} {p.description}")
        return 0
    profile = profiles.get(args.profile)
    if profile is None:
        ap.error(f"no profile named {args.profile!r} (see --list)")
    if profile.runner:
        # a profile with its own runner sets its own repetitions
        owned = [t for t in raw if t.split("=", 1)[0] in ("--reps", "--trials", "--components")]
        if owned:
            ap.error(f"{owned[0].split('=', 1)[0]} does not apply: profile {profile.name!r} runs under "
                     f"{profile.runner}, which sets its own repetitions")
        return subprocess.call([sys.executable, "-m", profile.runner, *[t for t in raw if t != args.profile]],
                               cwd=str(PKG))
    if not profile.runnable:
        ap.error(f"profile {profile.name!r} declares acceptance targets only; it has nothing to run")
    if args.reps < 1 or (args.trials is not None and args.trials < 1):
        ap.error("--reps and --trials take positive counts")
    components = [c.strip() for c in args.components.split(",") if c.strip()]
    try:
        cases = cases_of(profile, components)
    except ValueError as e:
        ap.error(str(e))
    trials = args.trials or REPEATED_PROFILES.get(profile.name, 1)
    print(f"profile {profile.name}: {len(cases)} cases x {trials} trials, {args.reps} reps each")
    if args.dry_run:
        for c in cases:
            print(c.shell(args.exe, args.reps))
        return 0
    if not Path(args.exe).is_file():
        ap.error(f"{args.exe} does not exist (build it: make -C ascend-fft_amd)")
    rows = run_cases(cases, trials, args.reps, args.exe, say=lambda s: print(s, flush=True))
    out = Path(args.out) if args.out else default_out(profile.name)
    s = write_outputs(out, profile.name, components, cases, trials, rows)
    print(f"{s['pass']}/{s['executions']} pass, {s['wrong']} wrong, {s['abnormal']} abnormal, "
          f"{s['not_run']} not run -> {out}")
    return 0 if s["result"] == "pass" else 1


if __name__ == "__main__":
    sys.exit(main())
