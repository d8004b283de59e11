"""CPU: the profile runner (bfly.suite) and the budgeted soak runner
(bfly.soak) -- case expansion, commands, outcome classification, the stop
at the first abnormal end, budgeting -- driven against bfly_check built on
the C ABI's CPU stub (which computes nothing, so every answer is wrong but
every run ends normally) and against stand-in programs that crash or hang
on the host."""
import csv
import json
import math
import os
import stat

import pytest

from bfly import _lib, soak, suite

PROFILES = suite.load_profiles()


def test_profiles_and_schema(tmp_path):
    assert {"current", "smoke", "regression", "publication", "long", "stress", "fft2d"} <= set(PROFILES)
    assert PROFILES["stress"].runner == "bfly.soak" and not PROFILES["fft2d"].runnable
    p = tmp_path / "m.json"
    p.write_text(json.dumps({"schema_version": 2, "profiles": {}}))
    with pytest.raises(ValueError):
        suite.load_profiles(p)


def test_case_expansion():
    cur = suite.cases_of(PROFILES["current"], ("c2c",))
    assert len(cur) == 49 and {c.input for c in cur} == {"sin"}
    groups = {}
    for c in suite.cases_of(PROFILES["smoke"]):
        groups[c.group] = groups.get(c.group, 0) + 1
    # 6 lengths x 5 batches; 6 real lengths x 4 batches; 3 cases x 7 patterns
    assert groups == {"c2c": 30, "real": 24, "numeric": 21}
    assert all(c.input == "half-spectrum" for c in suite.cases_of(PROFILES["smoke"]) if c.kind == "c2r")
    # random-seeded takes every listed seed, the other patterns seed 0 only
    reg = suite.cases_of(PROFILES["regression"], ("numeric",))
    seeded = {(c.n, c.batch, c.seed) for c in reg if c.input == "random-seeded"}
    assert len(seeded) == 6 * 3 and all(c.seed == 0 for c in reg if c.input != "random-seeded")
    apps = suite.cases_of(PROFILES["regression"], ("applications",))
    assert {(c.kind, c.input) for c in apps} == {("c2c", "ofdm"), ("c2c", "radar"), ("c2c", "dl"), ("r2c", "sin")}
    # equal-total-point shapes add batch = total / n, once each
    shapes = suite.c2c_shapes(PROFILES["long"].spec)
    assert (65536, 16) in shapes and (8192, 2048) in shapes and (16384, 64) in shapes
    assert len(shapes) == len(set(shapes))
    with pytest.raises(ValueError):
        suite.cases_of(PROFILES["smoke"], ("c2c", "fft2d"))


def test_commands():
    case = suite.Case("numeric", "c2c", 1024, 3, "random-seeded", 42)
    assert case.env() == {"AB_DIR": "c2c", "AB_INPUT": "random-seeded", "AB_SEED": "42"}
    assert case.argv("x", 5)[1:] == ["1024", "3", "5"]
    assert suite.Case("real", "c2r", 64, 1, "half-spectrum").env() == {"AB_DIR": "c2r"}


def test_outcome_classification():
    ok = "n=1 batch=1 launches=1 maxRel=2.000e-07 (worst idx 0)\nPASS\n"
    assert suite.classify(0, ok).status == "pass"
    assert suite.classify(0, "maxRel=2.000e-04\nPASS\n").status == "wrong"       # above the gate
    assert suite.classify(1, "maxRel=2.000e-07\nFAIL\n").status == "wrong"
    assert suite.classify(0, "maxRel=nan\nPASS\n").status == "wrong"
    assert suite.classify(0, "PASS\n").status == "wrong"                          # no error figure
    # a fault, a kill or a hang is not an answer
    assert suite.classify(-11, ok).status == "abnormal"
    assert suite.classify(0, ok, timed_out=True).status == "abnormal"
    assert suite.classify(134, "device init ok\n").status == "abnormal"           # no verdict line
    assert suite.classify(1, "device allocation failed\n").status == "abnormal"


def _fake(tmp_path, name, body):
    """A stand-in for bfly_check: a shell script (host only)."""
    p = tmp_path / name
    p.write_text("#!/bin/sh\n" + body + "\n")
    p.chmod(p.stat().st_mode | stat.S_IEXEC)
    return p


def test_first_abnormal_end_stops_the_run(tmp_path):
    # the stand-in segfaults on n = 256 and answers wrongly elsewhere
    exe = _fake(tmp_path, "check", 'if [ "$1" = 256 ]; then kill -SEGV $$; fi\n'
                                   'echo "n=$1 batch=$2 launches=1 maxRel=5.0e-01"; echo FAIL; exit 1')
    cases = [suite.Case("c2c", "c2c", n, 1, "sin") for n in (64, 128, 256, 512, 1024)]
    rows = suite.run_cases(cases, trials=2, reps=1, exe=exe, say=lambda s: None)
    assert [r["status"] for r in rows] == ["wrong", "wrong", "abnormal"] + ["not-run"] * 7
    assert rows[2]["note"].startswith("ended by signal 11")
    assert "n=256" in rows[5]["note"]
    s = suite.write_outputs(tmp_path / "o", "x", ["c2c"], cases, 2, rows)
    assert (s["wrong"], s["abnormal"], s["not_run"], s["result"]) == (2, 1, 7, "fail")


def test_time_limit_ends_a_case(tmp_path):
    exe = _fake(tmp_path, "hang", "sleep 30")
    o = suite.execute(suite.Case("c2c", "c2c", 64, 1, "sin"), 1, exe, limit_s=0.5)
    assert o.status == "abnormal" and "time limit" in o.note and o.wall_ms < 10000


def test_profile_run_on_the_stub(stub_cli, tmp_path, capsys):
    rc = suite.main(["smoke", "--components", "real", "--exe", str(stub_cli), "--out", str(tmp_path)])
    assert rc == 1
    rows = list(csv.DictReader((tmp_path / "cases.csv").open()))
    assert len(rows) == 24 and all(r["status"] == "wrong" for r in rows)
    s = json.loads((tmp_path / "summary.json").read_text())
    assert s["result"] == "fail" and s["wrong"] == 24 and s["executions"] == 24 and s["abnormal"] == 0
    assert suite.main(["current", "--dry-run", "--exe", str(stub_cli)]) == 0
    out = capsys.readouterr().out
    assert "profile current: 52 cases x 1 trials, 1 reps each" in out
    assert "AB_DIR=c2c AB_INPUT=sin" in out
    for bad in (["fft2d"], ["stress", "--reps", "3"], ["smoke", "--components", "c2c,fft2d"], ["nope"]):
        with pytest.raises(SystemExit):
            suite.main(bad)


def test_soak_budget_rules():
    b = soak.Budget(64 << 30, 16 << 30, 16 << 30)
    table = lambda n: 4 * _lib.stub().bfly_table_floats(_lib.KIND_C2C, n)
    fp = soak.Footprint.of("c2c", 1024, 128)
    assert b.refusal(fp) == "" and fp.device == 2 * 8 * 1024 * 128 + table(1024)
    # the four-step chain's workspace is counted; one workgroup per transform has none
    assert soak.Footprint.of("c2c", 65536, 2).device == 3 * 8 * 65536 * 2 + table(65536)
    assert soak.Footprint.of("c2c", 8192, 300).device == 2 * 8 * 8192 * 300 + table(8192)
    # device reserve: max(4 GiB, 5 %) below the free memory (16 GiB here)
    assert b.refusal(soak.Footprint.of("c2c", 1024, (11 << 30) // 16384)) == ""
    assert b.refusal(soak.Footprint.of("c2c", 1024, (13 << 30) // 16384)).startswith("device:")
    # 32-bit float offsets
    huge = soak.Budget(64 << 30, 1 << 40, 1 << 40)
    assert "32 bits" in huge.refusal(soak.Footprint.of("c2c", 64, (1 << 32) // 128))
    # host reserve
    assert soak.Budget(8 << 30, 1 << 40, 1 << 40).refusal(soak.Footprint.of("r2c", 8192, 1 << 18)).startswith("host:")
    jobs = soak.plan_jobs(PROFILES["stress"].spec, b)
    assert len(jobs) == 4 * 8
    soaks = [j for j in jobs if j.stage == "soak"]
    assert len(soaks) == 4 and all(j.case.n == 1024 and j.reps == 10000 for j in soaks)
    assert soak.batch_for_tier("c2r", 64, 1 << 20) == (1 << 20) // (4 * 66)
    assert soak.rss_growth([(i, 100) for i in range(5)]) is None
    samples = [(i, 100 << 20) for i in range(20)] + [(20 + i, 300 << 20) for i in range(4)]
    assert soak.rss_growth(samples) == 200 << 20 > soak.rss_allowance(300 << 20)


def test_soak_on_the_stub_reports_every_wrong_answer(stub_cli, tmp_path):
    rc = soak.main(["--lib", "stub", "--exe", str(stub_cli), "--tiers", "1048576", "--soak-reps", "50",
                    "--cycles", "2", "--oom-probe", "--device", "0", "--out", str(tmp_path)])
    assert rc == 1
    rows = list(csv.DictReader((tmp_path / "stress.csv").open()))
    ran = [r for r in rows if r["stage"] in ("case", "soak")]
    assert len(ran) == 8 and all(r["status"] == "wrong" for r in ran)
    probe = [r for r in rows if r["stage"] == "oom_probe"][0]
    # the stub refuses the oversize request but computes no transform
    assert probe["status"] == "wrong" and "wrong spectrum" in probe["note"]
    s = json.loads((tmp_path / "summary.json").read_text())
    assert s["result"] == "fail" and s["lifecycle_cycles"] == 0 and s["soak_reps"] == 50
    with pytest.raises(SystemExit):
        soak.main(["--lib", "stub", "--oom-probe"])           # needs an explicit device


def test_soak_stops_at_an_abnormal_end(tmp_path):
    # passes everywhere except n = 64, which aborts: nothing after it runs,
    # neither the lifecycle nor the probe
    exe = _fake(tmp_path, "check", 'if [ "$1" = 64 ]; then kill -ABRT $$; fi\n'
                                   'echo "n=$1 batch=$2 launches=1 maxRel=1.0e-07"; echo PASS')
    spec = {"c2c": {"ns": [4096, 64, 1024]}, "real": {"r2c_ns": [128], "c2r_ns": []},
            "input_buffer_bytes": [1 << 20], "repeat_execution": 5}
    jobs = soak.plan_jobs(spec, soak.Budget(64 << 30, 1 << 40, 1 << 40))
    done = soak.run(jobs, exe, cycles=3, probe_device=0, lib=_lib.stub(), say=lambda s: None)
    status = {(j.stage, j.case.n if j.case else None): j.outcome.status for j in done}
    assert status[("case", 4096)] == "pass" and status[("case", 64)] == "abnormal"
    assert status[("soak", 1024)] == status[("case", 128)] == "not-run"
    assert status[("lifecycle", 4096)] == status[("oom_probe", None)] == "not-run"


def test_stub_reports_its_memory_and_refuses_oversize():
    import ctypes as C
    s = _lib.stub()
    free, total = soak.device_memory(s)
    assert free == total == 16 << 30
    p = C.c_void_p()
    assert s.bfly_rt_malloc(C.byref(p), 2 * total) == -3 and not p.value
