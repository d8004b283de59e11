"""CPU: the profile runner (bfly.matrix) and the budgeted stress runner
(bfly.stress) -- case expansion, commands, verdict parsing, budgeting and
their failure reporting, driven against the bfly_check program built on
the C ABI's CPU stub (which computes nothing, so every run must be
reported as failed).  Mirrors the reference's scripts/run_test_profile.py
and scripts/run_stress.py behaviour."""
import csv
import json

import pytest

from bfly import _lib, matrix, stress

CONFIG = matrix.load_config()
PROFILES = CONFIG["profiles"]


def test_profiles_and_schema(tmp_path):
    assert {"current", "smoke", "regression", "publication", "long", "stress", "fft2d"} <= set(PROFILES)
    p = tmp_path / "m.json"
    p.write_text(json.dumps({"schema_version": 2, "profiles": {}}))
    with pytest.raises(ValueError):
        matrix.load_config(p)


def test_case_expansion():
    cur = matrix.expand_cases(PROFILES["current"], ("c2c",))
    assert len(cur) == 49 and {c["input"] for c in cur} == {"sin"}
    smoke = matrix.expand_cases(PROFILES["smoke"])
    cats = {}
    for c in smoke:
        cats[c["category"]] = cats.get(c["category"], 0) + 1
    # 6 lengths x 5 batches; 6 real lengths x 4 batches; 3 cases x 7 patterns
    assert cats == {"c2c": 30, "real": 24, "numeric": 21}
    assert all(c["input"] == "half-spectrum" for c in smoke if c["direction"] == "c2r")
    # random-seeded takes every listed seed, the other patterns seed 0 only
    reg = matrix.expand_cases(PROFILES["regression"], ("numeric",))
    seeds = {(c["n"], c["batch"], c["seed"]) for c in reg if c["input"] == "random-seeded"}
    assert len(seeds) == 6 * 3 and all(c["seed"] == 0 for c in reg if c["input"] != "random-seeded")
    apps = matrix.expand_cases(PROFILES["regression"], ("applications",))
    assert {(c["direction"], c["input"]) for c in apps} == {("c2c", "ofdm"), ("c2c", "radar"), ("c2c", "dl"),
                                                             ("r2c", "sin")}
    # fixed total points add batch = total / n, deduplicated with the grid
    pts = matrix.c2c_points(PROFILES["long"])
    assert (65536, 16) in pts and (8192, 2048) in pts and (16384, 64) in pts
    assert len(pts) == len(set(pts))


def test_commands_and_verdicts(stub_cli):
    case = {"category": "numeric", "direction": "c2c", "n": 1024, "batch": 3, "input": "random-seeded", "seed": 42}
    env, argv = matrix.command_for(case, 5, stub_cli)
    assert env == {"AB_DIR": "c2c", "AB_INPUT": "random-seeded", "AB_SEED": "42"}
    assert argv[1:] == ["1024", "3", "5"]
    half = dict(case, direction="c2r", input="half-spectrum", seed=0)
    assert matrix.command_for(half, 1)[0] == {"AB_DIR": "c2r"}
    assert matrix.verdict_of(0, "n=1 maxRel=2.000e-07 (worst idx 0)\nPASS\n") == (2e-7, True)
    assert matrix.verdict_of(0, "maxRel=2.000e-04\nPASS\n")[1] is False       # above the gate
    assert matrix.verdict_of(1, "maxRel=2.000e-07\nFAIL\n")[1] is False
    assert matrix.verdict_of(0, "maxRel=nan\nPASS\n")[1] is False
    assert matrix.verdict_of(0, "PASS\n")[1] is False                        # no error figure


def test_profile_run_reports_failures(stub_cli, tmp_path, capsys):
    rc = matrix.main(["smoke", "--components", "real", "--exe", str(stub_cli), "--out", str(tmp_path)])
    assert rc == 1
    rows = list(csv.DictReader((tmp_path / "cases.csv").open()))
    assert len(rows) == 24 and all(r["correct"] == "False" for r in rows)
    s = json.loads((tmp_path / "summary.json").read_text())
    assert s["status"] == "fail" and s["failed"] == 24 and s["executions"] == 24
    assert matrix.main(["current", "--dry-run", "--exe", str(stub_cli)]) == 0
    out = capsys.readouterr().out
    assert "profile=current cases=52 trials=1 reps=1" in out
    assert "AB_DIR=c2c AB_INPUT=sin" in out
    with pytest.raises(SystemExit):
        matrix.main(["fft2d"])                    # acceptance target only
    with pytest.raises(SystemExit):
        matrix.main(["stress", "--reps", "3"])    # the dedicated runner owns its budget
    with pytest.raises(SystemExit):
        matrix.main(["smoke", "--components", "c2c,fft2d"])


def test_stress_budget_rules():
    host, free, total = 64 << 30, 16 << 30, 16 << 30
    b = stress.budget_case("c2c", 1024, 128, host, free, total)
    assert b["decision"] == "run" and b["device_needed"] == 2 * 8 * 1024 * 128 + stress.table_bytes("c2c", 1024)
    # the four-step chain's workspace is counted
    assert stress.device_bytes_exact("c2c", 65536, 2) == 3 * 8 * 65536 * 2 + stress.table_bytes("c2c", 65536)
    assert stress.device_bytes_exact("c2c", 8192, 300) == 2 * 8 * 8192 * 300 + stress.table_bytes("c2c", 8192)
    # device headroom: max(4 GiB, 5 %) below the free memory (16 GiB here)
    assert stress.budget_case("c2c", 1024, (11 << 30) // 16384, host, free, total)["decision"] == "run"
    assert stress.budget_case("c2c", 1024, (13 << 30) // 16384, host, free, total)["decision"] == "skip"
    # 32-bit float offsets
    big = stress.budget_case("c2c", 64, (1 << 32) // 128, host, 1 << 40, 1 << 40)
    assert big["decision"] == "skip" and "offset" in big["reason"]
    # host margin
    assert "host needs" in stress.budget_case("r2c", 8192, 1 << 18, 8 << 30, 1 << 40, 1 << 40)["reason"]
    rows = stress.plan(PROFILES["stress"], host, free, total)
    assert len(rows) == 4 * 8
    soaks = [r for r in rows if r["stage"] == "soak"]
    assert len(soaks) == 4 and all(r["n"] == 1024 and r["reps"] == 10000 for r in soaks)
    assert stress.derive_batch(64, 1 << 20, "c2r") == (1 << 20) // (4 * 66)
    assert stress.rss_growth_bytes([(i, 100) for i in range(5)]) is None
    samples = [(i, 100 << 20) for i in range(20)] + [(20 + i, 300 << 20) for i in range(4)]
    assert stress.rss_growth_bytes(samples) == 200 << 20 > stress.rss_limit_bytes(300 << 20)


def test_stress_run_on_the_stub_reports_every_failure(stub_cli, tmp_path):
    rc = stress.main(["--lib", "stub", "--exe", str(stub_cli), "--tiers", "1048576", "--soak-reps", "50",
                      "--lifecycle-cycles", "2", "--oom-probe", "--device", "0", "--out", str(tmp_path)])
    assert rc == 1
    rows = list(csv.DictReader((tmp_path / "stress.csv").open()))
    ran = [r for r in rows if r["decision"] == "run"]
    assert len(ran) == 9 and all(r["correct"] == "False" for r in ran)
    probe = [r for r in rows if r["stage"] == "oom_probe"][0]
    # the stub refuses the oversize allocation but computes no transform
    assert "wrong transform" in probe["failure"]
    s = json.loads((tmp_path / "summary.json").read_text())
    assert s["status"] == "fail" and s["lifecycle_cycles"] == 0 and s["soak_reps"] == 50
    with pytest.raises(SystemExit):
        stress.main(["--lib", "stub", "--oom-probe"])           # needs an explicit device


def test_stub_reports_its_memory_and_refuses_oversize():
    import ctypes as C
    s = _lib.stub()
    free, total = stress.read_device_memory(s)
    assert free == total == 16 << 30
    p = C.c_void_p()
    assert s.bfly_rt_malloc(C.byref(p), 2 * total) == -3 and not p.value
