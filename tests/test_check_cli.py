"""CPU: the bfly_check program (ascend-fft_amd/csrc/check/) against the C
ABI's CPU stub, and the report reader bfly.report.

The stub runs every guard but computes nothing, which makes it the planted
mistake for the program's own checks: every transform must FAIL and a
sequence of different inputs must be flagged STALE-OUTPUT.  The generated
inputs (dumped with AB_DUMP) must equal the oracle's restatement of the
reference's fft_check generators bit for bit.  Argument handling follows
src/host/fft_check.cpp:213-300 (exit 2 on a bad argument).
"""
import os
import subprocess
from pathlib import Path

import numpy as np
import pytest

import oracle
from bfly import report

ROOT = Path(__file__).resolve().parents[1]
PKG = ROOT / "ascend-fft_amd"


@pytest.fixture(scope="module")
def cli(stub_cli):
    return stub_cli


def run(exe, *args, **envs):
    env = {k: v for k, v in os.environ.items() if not k.startswith("AB_")}
    env.update({k: str(v) for k, v in envs.items()})
    return subprocess.run([str(exe), *map(str, args)], capture_output=True, text=True, env=env,
                          timeout=120)


@pytest.mark.parametrize("args,envs,msg", [
    ((), {}, "usage"),
    ((100, 2), {}, "n must be power of two"),
    ((4, 2), {}, "n must be power of two"),
    ((64, 2), {"AB_DIR": "fft"}, "AB_DIR must be"),
    ((64, 2), {"AB_INPUT": "square"}, "unsupported AB_INPUT mode"),
    ((64, 2), {"AB_PLANE_K": "12"}, "AB_PLANE_K must be one of 8|16|32"),
    ((256, 2), {"AB_DIR": "r2c", "AB_INPUT_SEQ": "sin,zero"}, "only supported for c2c"),
    ((256, 2), {"AB_INPUT_SEQ": "sin"}, "at least two inputs"),
    ((131072, 1), {}, "outside the library's envelope"),
    ((16384, 1), {"AB_DIR": "c2r"}, "outside the library's envelope"),
    ((8192, 2), {"AB_BOUNDARY": "host"}, "AB_BOUNDARY must be device"),
    ((8192, 2), {"AB_LONG_BOUNDARY_IMPL": "separate"}, "AB_LONG_BOUNDARY_IMPL must be fused"),
    ((256, 2), {"AB_DIR": "r2c", "AB_SGN": "+1"}, "fixed sign"),
    ((256, 2), {"AB_VARIANT": "fastest"}, "unknown AB_VARIANT"),
    ((256, 2), {"AB_DIR": "r2c", "AB_VARIANT": "stream2"}, "does not apply to r2c"),
    ((256, 2), {"AB_INPUT_FILE": "/nonexistent/in.bin"}, "AB_INPUT_FILE read failed"),
])
def test_bad_arguments_exit_2(cli, args, envs, msg):
    r = run(cli, *args, **envs)
    assert r.returncode == 2, r.stdout + r.stderr
    assert msg in r.stdout


def _dump(cli, tmp_path, n, b, **envs):
    prefix = tmp_path / "d"
    r = run(cli, n, b, 1, AB_DUMP=prefix, **envs)
    # the stub computes nothing: the verdict must be FAIL
    assert r.returncode == 1 and r.stdout.rstrip().endswith("FAIL"), r.stdout + r.stderr
    return np.fromfile(f"{prefix}.in.bin", np.float32)


@pytest.mark.parametrize("mode,alias", [
    ("zero", None), ("impulse", None), ("constant", None), ("tone", "single-tone"),
    ("alternating", "nyquist-alternating"), ("random-seeded", "random"),
    ("high-dynamic-range", "dynamic"), ("near-cancellation", None)])
def test_numerical_inputs_match_the_oracle(cli, tmp_path, mode, alias):
    n, b = 64, 9
    want = oracle.numerical_input(n, b, mode, seed=0x1234).reshape(-1)
    for name in (mode, alias) if alias else (mode,):
        got = _dump(cli, tmp_path, n, b, AB_INPUT=name, AB_SEED="0x1234")
        np.testing.assert_array_equal(got.view(np.uint32), want.view(np.uint32))


@pytest.mark.parametrize("mode", ["ofdm", "radar", "dl"])
def test_application_inputs_match_the_oracle(cli, tmp_path, mode):
    got = _dump(cli, tmp_path, 256, 5, AB_INPUT=mode)
    np.testing.assert_array_equal(got.view(np.uint32), oracle.app_input(256, 5, mode).reshape(-1).view(np.uint32))


def test_sin_inputs_and_real_directions_match_the_oracle(cli, tmp_path):
    got = _dump(cli, tmp_path, 128, 3)
    np.testing.assert_array_equal(got, oracle.sin_input(2 * 128 * 3))
    got = _dump(cli, tmp_path, 128, 3, AB_DIR="r2c")
    np.testing.assert_array_equal(got, oracle.sin_input(128 * 3))
    # R2C of a complex mode takes the real parts
    got = _dump(cli, tmp_path, 128, 3, AB_DIR="r2c", AB_INPUT="random-seeded", AB_SEED=5)
    np.testing.assert_array_equal(got, oracle.numerical_input(128, 3, "random-seeded", seed=5)[:, 0::2].reshape(-1))
    # C2R always takes the seeded half spectrum, rows of n+2 floats
    got = _dump(cli, tmp_path, 256, 4, AB_DIR="c2r", AB_INPUT="impulse")
    np.testing.assert_array_equal(got, oracle.half_input(256, 4).reshape(-1))


def test_input_file_replaces_the_generated_input(cli, tmp_path):
    x = oracle.uniform_input(3, (2, 2 * 64))
    f = tmp_path / "x.bin"
    x.tofile(f)
    got = _dump(cli, tmp_path, 64, 2, AB_INPUT_FILE=f)
    np.testing.assert_array_equal(got, x.reshape(-1))


def test_stale_outputs_are_flagged(cli):
    # outputs that do not follow their inputs (the stub writes nothing)
    for n in (1024, 8192, 65536):
        r = run(cli, n, 2, 1, AB_INPUT_SEQ="sin,random-seeded,sin")
        assert r.returncode == 1, r.stdout
        rep = report.Report.parse(r.stdout)
        assert [s.stale for s in rep.seq] == [False, True, True], r.stdout
        assert "stale-check on -> FAIL" in r.stdout
        # identical consecutive inputs are not stale
        r = run(cli, n, 2, 1, AB_INPUT_SEQ="zero,zero")
        assert [s.stale for s in report.Report.parse(r.stdout).seq] == [False, False]


def test_report_lines_follow_the_vocabulary(cli):
    r = run(cli, 1024, 4, 3)
    rep = report.Report.parse(r.stdout)
    assert rep.verdict == "FAIL" and rep.launches == 1 and rep.chain is None
    f = rep.scopes
    assert f.reps == 3 and f.h2d is None and f.d2h is None and rep.segments is None
    assert rep.mode == "resident"
    for n, b, chain, segs in ((8192, 300, "single", {"fft1"}), (16384, 2, "fused2", {"fft1", "fft2"}),
                              (65536, 1, "fused2", {"fft1", "fft2"})):
        rep = report.Report.parse(run(cli, n, b, 2).stdout)
        assert rep.chain == chain and set(rep.segments) == segs, (n, b)
        assert rep.scopes.h2d is not None and rep.scopes.d2h is not None
        assert rep.mode == "staged"
    rep = report.Report.parse(run(cli, 8192, 3, 2, AB_VARIANT="four_step").stdout)
    assert rep.chain == "fused2" and rep.launches == 2
    # the lowering descriptor of the chain (include/butterfly/descriptors.hpp) and its split
    out = run(cli, 65536, 2, 1, AB_DESC=1).stdout
    assert "lowering: launches=2 gm_boundaries=1 modeled_payload_gm=4194304 impl=fused2" in out
    assert "kinds=row_fft|row_fft" in out
    assert "long backend: n=65536 split N1=256 N2=256 rows1=512 rows2=512 gm_bytes=3145728" in out
    out = run(cli, 16384, 300, 1, AB_DESC=1).stdout
    assert "lowering: launches=1 gm_boundaries=0" in out and "resident=block-resident on_chip=1" in out
    assert "long backend: n=16384 split N1=16384 N2=1 rows1=300 rows2=0" in out
    assert "lowering:" not in run(cli, 65536, 2, 1).stdout          # only with AB_DESC
    assert run(cli, 64, 2, 1, AB_VARIANT="split_mfma").returncode == 2   # variant does not apply
    r = run(cli, 4096, 2, 2, AB_E2E=3, AB_E2E_MODE="xfer")
    assert "E2E n=4096 batch=2 reps=3" in r.stdout and "mode=xfer host=pinned" in r.stdout
    assert "E2E transfers: in=1 out=1 boundary=0 per_execution" in r.stdout
    # the pipelined host-to-host call checks its own output: the stub computes
    # nothing, so the check fails and the program stops with status 1
    r = run(cli, 1024, 5, 1, AB_E2E=2, AB_E2E_MODE="pipe", AB_E2E_CHUNK=2)
    assert "E2E n=1024 batch=5 reps=2" in r.stdout and "mode=pipe host=pinned" in r.stdout
    assert "E2E pipe maxRel=" in r.stdout and "FAIL" in r.stdout and r.returncode == 1


# ---------------------------------------------------------------- reader
STAGED = ("scopes: plan_setup=812.5 us first_use=950.0 us host_end_to_end mean=4100.0 min=3900.0 us "
          "h2d=1500.0 us device_chain=210.0 us d2h=1450.0 us reps=5")
RESIDENT = ("scopes: plan_setup=300.0 us first_use=400.0 us host_end_to_end mean=60.0 min=55.0 us "
            "h2d=NA device_chain=47.5 us d2h=NA reps=20")


def test_token_reader():
    assert report.read_pairs("a=1 us lab b=2 c=NA us d=3") == {"a": "1", "lab.b": "2", "lab.c": None, "d": "3"}
    for bad in ("a=1 a=2", "x y a=1", "=3", "a="):
        with pytest.raises(report.ReportError):
            report.read_pairs(bad)


def test_scopes_line_fields_and_na():
    f = report.Scopes.parse(STAGED)
    assert (f.h2d, f.device_chain, f.d2h, f.e2e_mean, f.e2e_min, f.reps) == (1500.0, 210.0, 1450.0, 4100.0, 3900.0, 5)
    assert f.mode == "staged"
    s = report.Scopes.parse(RESIDENT)
    assert s.h2d is None and s.d2h is None and s.device_chain == 47.5 and s.mode == "resident"
    for bad in (STAGED.replace(" d2h=1450.0 us", ""), "PASS", STAGED.replace("reps=5", "reps=five"),
                "scopes: plan_setup=1.0 us first_use=2.0 us host_end_to_end mean=3.0 min=2.0 us "
                "device_only=1.0 us reps=1"):
        with pytest.raises(report.ReportError):
            report.Scopes.parse(bad)


def test_scope_rules():
    staged, resident = report.Scopes.parse(STAGED), report.Scopes.parse(RESIDENT)
    assert report.scope_problems(resident) == [] and report.scope_problems(staged) == []
    # a resident run must not time copies; a staged one must
    assert any("want NA" in p for p in report.scope_problems(staged, "resident"))
    assert any("must time both copies" in p for p in report.scope_problems(resident, "staged"))
    # copies counted inside the chain overrun the wall clock
    bad = report.Scopes.parse(STAGED.replace("device_chain=210.0", "device_chain=9000.0"))
    assert any("copies were counted inside device_chain" in p for p in report.scope_problems(bad))
    # the tolerance is the caller's
    assert report.scope_problems(bad, tol=report.Tolerance(wall_us=1e5)) == []
    assert report.scope_problems(resident, "sideways")[0].startswith("mode 'sideways'")


def test_segments_and_chain_forms():
    assert report.parse_segments("segments: NA") is None
    seg = report.parse_segments("segments: fft1=120.0 us fft2=90.0 us")
    assert seg == {"fft1": 120.0, "fft2": 90.0}
    f = report.Scopes.parse(STAGED)
    assert report.segment_problems(seg, f, "fused2") == []
    assert report.segment_problems({"fft1": 210.0}, f, "single") == []
    assert any("add up to" in p for p in report.segment_problems({"fft1": 500.0, "fft2": 90.0}, f, "fused2"))
    assert any("is not timed" in p for p in report.segment_problems({"fft1": 210.0}, f, "fused2"))
    assert any("does not belong" in p for p in report.segment_problems(seg, f, "single"))
    resident = report.Scopes.parse(RESIDENT)
    assert report.segment_problems(None, resident, None) == []
    assert report.segment_problems(seg, resident, None) != []
    six = {k: 35.0 for k in report.LAUNCH_NAMES}
    assert report.segment_problems(six, f, "separate") == []
    with pytest.raises(report.ReportError):
        report.parse_segments("segments: fft3=1.0 us")
    for form in report.CHAINS:
        assert report.parse_chain(f"boundary_impl: {form}") == form
    with pytest.raises(report.ReportError):
        report.parse_chain("boundary_impl: bespoke")


def test_whole_report():
    text = "\n".join([STAGED, "boundary_impl: fused2", "segments: fft1=120.0 us fft2=90.0 us",
                      "n=65536 batch=2 launches=2  maxAbs=1e-3 maxRel=2.0e-07 (worst idx 5)",
                      "bfly_c2c: 4100.0 us/call over 5 reps (min 3900.0 us) med 4100.0 us", "PASS"])
    rep = report.Report.parse(text)
    assert rep.problems() == [] and rep.verdict == "PASS" and rep.launches == 2 and rep.us_per_call == 4100.0
    assert report.Report.parse(text.replace("PASS", "")).problems() == ["no PASS / FAIL line"]
