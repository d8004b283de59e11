"""Run a predeclared correctness profile of config/test_matrix.json through
the bfly_check program, one process per case.

The counterpart of the reference's scripts/run_test_profile.py:1-255: the
same profile schema, case expansion (c2c grid + fixed-total-point shapes,
real rows, numerical patterns with their seeds, application shapes), PASS
criterion (exit 0, a `PASS` line and 0 <= maxRel <= 1e-4), dedicated-runner
hand-off (the stress profile goes to bfly.stress) and outputs (cases.csv and
summary.json in the run directory).  Timing publication stays with bench.py.

    python -m bfly.matrix smoke [--reps K] [--trials T] [--components c2c,real]
                                [--dry-run] [--list] [--out DIR] [--exe PATH]
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
DEFAULT_CONFIG = PKG / "config" / "test_matrix.json"
DEFAULT_EXE = PKG / "lib" / "bfly_check"
DEFAULT_COMPONENTS = ("c2c", "real", "numeric", "applications")
THRESHOLD = 1e-4
FIELDS = ("category", "direction", "n", "batch", "input", "seed", "trial",
          "returncode", "max_rel", "wall_ms", "device_us", "correct", "failure", "output_tail")
# application ids of the profile schema -> bfly_check input modes
APP_MODES = {"ofdm": "ofdm", "radar-range": "radar", "dl-frequency-layer": "dl", "stft": "sin"}


def load_config(path=DEFAULT_CONFIG):
    doc = json.loads(Path(path).read_text(encoding="utf-8"))
    if doc.get("schema_version") != 1 or not isinstance(doc.get("profiles"), dict):
        raise ValueError("unsupported test-matrix schema")
    return doc


def c2c_points(profile):
    """(n, batch) pairs of the c2c grid plus the fixed-total-point shapes
    (batch = total / n for every n dividing the total), deduplicated."""
    spec = profile.get("c2c", {})
    ns = [int(n) for n in spec.get("ns", [])]
    pts = {(n, int(b)) for n in ns for b in spec.get("batches", [])}
    for total in spec.get("fixed_total_points", []):
        pts.update((n, int(total) // n) for n in ns if int(total) % n == 0)
    return sorted(pts)


def _case(category, direction, n, batch, mode, seed=0):
    return {"category": category, "direction": direction, "n": int(n), "batch": int(batch),
            "input": mode, "seed": int(seed)}


def expand_cases(profile, components=DEFAULT_COMPONENTS):
    """Every case of a profile, sorted and unique on all six fields."""
    components = set(components)
    cases = []
    if "c2c" in components:
        cases += [_case("c2c", "c2c", n, b, "sin") for n, b in c2c_points(profile)]
    if "real" in components:
        real = profile.get("real", {})
        for direction, key, mode in (("r2c", "r2c_ns", "sin"), ("c2r", "c2r_ns", "half-spectrum")):
            cases += [_case("real", direction, n, b, mode)
                      for n in real.get(key, []) for b in real.get("batches", [])]
    if "numeric" in components:
        seeds = profile.get("random_seeds", [0])
        for n, b in profile.get("numeric_cases", []):
            for pattern in profile.get("numeric_patterns", []):
                for seed in (seeds if pattern == "random-seeded" else [0]):
                    cases.append(_case("numeric", "c2c", n, b, pattern, seed))
    if "applications" in components:
        for app in profile.get("applications", []):
            if app.get("transform") not in ("c2c", "r2c") or app.get("id") not in APP_MODES:
                continue
            cases += [_case("application", app["transform"], n, b, APP_MODES[app["id"]])
                      for n in app.get("ns", []) for b in app.get("batches", [])]
    unique = {tuple(c[k] for k in ("category", "direction", "n", "batch", "input", "seed")): c
              for c in cases}
    return [unique[k] for k in sorted(unique)]


def command_for(case, reps, exe=DEFAULT_EXE):
    """(environment additions, argv) of one case."""
    env = {"AB_DIR": case["direction"]}
    if case["input"] != "half-spectrum":
        env["AB_INPUT"] = case["input"]
    if case["seed"]:
        env["AB_SEED"] = str(case["seed"])
    return env, [str(exe), str(case["n"]), str(case["batch"]), str(reps)]


def shell_line(case, reps, exe=DEFAULT_EXE):
    env, argv = command_for(case, reps, exe)
    return " ".join(f"{k}={v}" for k, v in sorted(env.items())) + " " + " ".join(argv)


def verdict_of(returncode, output):
    """(max_rel, correct) of one bfly_check run: exit 0, a PASS line and a
    finite maxRel in [0, 1e-4]."""
    m = re.search(r"maxRel=([0-9.eE+-]+|nan|inf)", output)
    err = float(m.group(1)) if m else math.nan
    ok = (returncode == 0 and re.search(r"^PASS$", output, re.M) is not None
          and math.isfinite(err) and 0.0 <= err <= THRESHOLD)
    return err, ok


def run_case(case, reps, trial, exe=DEFAULT_EXE, timeout=600.0):
    extra, argv = command_for(case, reps, exe)
    env = {k: v for k, v in os.environ.items() if not k.startswith("AB_")}
    env.update(extra)
    t0 = time.perf_counter()
    try:
        r = subprocess.run(argv, env=env, capture_output=True, text=True, timeout=timeout)
    except subprocess.TimeoutExpired as exc:
        return dict(case, trial=trial, returncode=124, max_rel=math.nan,
                    wall_ms=(time.perf_counter() - t0) * 1e3, device_us=math.nan, correct=False,
                    failure=f"timeout after {timeout:.0f} s", output_tail=str(exc))
    wall_ms = (time.perf_counter() - t0) * 1e3
    out = r.stdout + r.stderr
    err, ok = verdict_of(r.returncode, out)
    m = re.search(r"device_chain=([0-9.]+) us", out)
    return dict(case, trial=trial, returncode=r.returncode, max_rel=err, wall_ms=wall_ms,
                device_us=float(m.group(1)) if m else math.nan, correct=ok,
                failure="" if ok else "exit status, PASS line or maxRel check failed",
                output_tail="" if ok else "\n".join(out.splitlines()[-12:]))


def write_results(out, profile, components, cases, trials, rows):
    out.mkdir(parents=True, exist_ok=True)
    with (out / "cases.csv").open("w", newline="", encoding="utf-8") as fh:
        w = csv.DictWriter(fh, fieldnames=FIELDS)
        w.writeheader()
        w.writerows(rows)
    failed = sum(1 for r in rows if not r["correct"])
    errs = [r["max_rel"] for r in rows if math.isfinite(r["max_rel"])]
    summary = {"schema_version": 1, "profile": profile, "components": list(components),
               "cases": len(cases), "trials": trials, "executions": len(rows),
               "passed": len(rows) - failed, "failed": failed, "correctness_threshold": THRESHOLD,
               "worst_max_rel": max(errs) if errs else None,
               "status": "pass" if not failed else "fail"}
    (out / "summary.json").write_text(json.dumps(summary, indent=2) + "\n", encoding="utf-8")
    return summary


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("profile", nargs="?", default="smoke")
    ap.add_argument("--config", default=str(DEFAULT_CONFIG))
    ap.add_argument("--exe", default=str(DEFAULT_EXE))
    ap.add_argument("--list", action="store_true")
    ap.add_argument("--components", default=",".join(DEFAULT_COMPONENTS))
    ap.add_argument("--trials", type=int, default=None)
    ap.add_argument("--reps", type=int, default=1)
    ap.add_argument("--dry-run", action="store_true")
    ap.add_argument("--out", default="")
    raw = list(sys.argv[1:] if argv is None else argv)
    args = ap.parse_args(raw)
    doc = load_config(args.config)
    profiles = doc["profiles"]
    if args.list:
        for name, p in profiles.items():
            print(f"{name:12} runnable={str(bool(p.get('runnable'))).lower():5} {p.get('description', '')}")
        return 0
    if args.profile not in profiles:
        ap.error(f"unknown profile {args.profile!r}; use --list")
    profile = profiles[args.profile]
    if profile.get("runner"):
        # a dedicated runner owns its repetition budget
        fwd = [t for t in raw if t != args.profile]
        for t in fwd:
            if t.split("=", 1)[0] in ("--reps", "--trials", "--components"):
                ap.error(f"profile {args.profile!r} owns its repetition budget; "
                         f"{t.split('=', 1)[0]} is managed by {profile['runner']}")
        return subprocess.call([sys.executable, "-m", profile["runner"], *fwd], cwd=str(PKG))
    if not profile.get("runnable"):
        ap.error(f"profile {args.profile!r} is an acceptance target, not runnable")
    components = tuple(c.strip() for c in args.components.split(",") if c.strip())
    unknown = set(components) - set(DEFAULT_COMPONENTS)
    if unknown:
        ap.error(f"unknown components: {','.join(sorted(unknown))}")
    if args.reps < 1 or (args.trials is not None and args.trials < 1):
        ap.error("reps and trials must be positive")
    trials = args.trials or (3 if args.profile in ("regression", "publication") else 1)
    cases = expand_cases(profile, components)
    print(f"profile={args.profile} cases={len(cases)} trials={trials} reps={args.reps}")
    if args.dry_run:
        for c in cases:
            print(shell_line(c, args.reps, args.exe))
        return 0
    if not Path(args.exe).is_file():
        ap.error(f"{args.exe} is missing (make -C ascend-fft_amd)")
    stamp = datetime.now(timezone.utc).strftime("%Y%m%dT%H%M%SZ")
    out = Path(args.out) if args.out else Path.cwd() / "bfly_runs" / f"{stamp}-{args.profile}"
    rows, total, i = [], len(cases) * trials, 0
    for trial in range(1, trials + 1):
        for c in cases:
            i += 1
            print(f"[{i}/{total}] {c['category']} {c['direction']} N={c['n']} B={c['batch']} "
                  f"input={c['input']} trial={trial}", flush=True)
            row = run_case(c, args.reps, trial, args.exe)
            rows.append(row)
            if not row["correct"]:
                print(row["output_tail"], file=sys.stderr)
    s = write_results(out, args.profile, components, cases, trials, rows)
    print(f"{s['passed']}/{s['executions']} PASS -> {out}")
    return 0 if not s["failed"] else 1


if __name__ == "__main__":
    sys.exit(main())
