"""GPU: the bfly_check program (the counterpart of the reference's
build/fft_check, src/host/fft_check.cpp:213-1315) on the MI355X.

Every run must PASS its own gate (maxRelScaled <= 1e-4 per row against the
library's FP64 reference), print timing lines that validate under
bfly.report (resident runs: no transfers; staged long chains: h2d / d2h measured and
segments adding up to device_chain), and its dumped output must meet the
FP64 oracle at relative L2 <= 1e-5 per row.  AB_INPUT_SEQ must re-run one
plan on changing inputs with no stale output.
"""
import os
import subprocess
from pathlib import Path

import numpy as np
import pytest

import oracle
from bfly import report
from conftest import gpu_available

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not gpu_available(), reason="needs an MI355X")]

ROOT = Path(__file__).resolve().parents[1]
EXE = ROOT / "ascend-fft_amd" / "lib" / "bfly_check"
REL_L2 = 1e-5


def run(*args, **envs):
    if not EXE.exists():
        pytest.fail("bfly_check is not built (make -C ascend-fft_amd)")
    env = {k: v for k, v in os.environ.items() if not k.startswith("AB_")}
    env.update({k: str(v) for k, v in envs.items()})
    print("bfly_check", *args, envs, flush=True)
    r = subprocess.run([str(EXE), *map(str, args)], capture_output=True, text=True, env=env, timeout=180)
    print(r.stdout, flush=True)
    return r


def _oracle_ref(kind, x, n, sign=-1):
    if kind == "c2c":
        return oracle.c2c_f64(x, n, sign)
    if kind == "r2c":
        return oracle.r2c_f64(x, n)
    return oracle.c2r_f64(x, n)


CASES = [
    # (n, batch, reps, env)
    (1024, 64, 5, {}),                                                  # the 1024 x 64 latency shape
    (4096, 16, 3, {"AB_INPUT": "random-seeded", "AB_SEED": 7}),
    (64, 7, 3, {"AB_INPUT": "impulse"}),
    (16, 5, 3, {"AB_INPUT": "tone"}),
    (2048, 140, 3, {"AB_INPUT": "ofdm"}),
    (1024, 256, 3, {"AB_INPUT": "radar"}),
    (4096, 32, 3, {"AB_INPUT": "dl"}),
    (512, 33, 3, {"AB_INPUT": "high-dynamic-range"}),
    (4096, 9, 3, {"AB_SGN": "+1", "AB_INPUT": "random-seeded"}),
    (4096, 64, 3, {"AB_VARIANT": "split_mfma", "AB_INPUT": "random-seeded"}),
    (4096, 8, 3, {"AB_DIR": "r2c"}),
    (8192, 5, 3, {"AB_DIR": "r2c", "AB_INPUT": "random-seeded"}),
    (128, 49, 3, {"AB_DIR": "r2c", "AB_INPUT": "near-cancellation"}),
    (1024, 5, 3, {"AB_DIR": "c2r"}),
    (8192, 3, 3, {"AB_DIR": "c2r"}),
    (8192, 300, 3, {}),                                                 # one launch (single)
    (16384, 4, 3, {"AB_INPUT": "random-seeded"}),                       # four-step below 256 rows
    (8192, 3, 3, {"AB_VARIANT": "four_step", "AB_INPUT": "radar"}),
    (32768, 3, 3, {"AB_INPUT": "random-seeded", "AB_SEED": 2}),
    (65536, 2, 3, {"AB_INPUT": "random-seeded", "AB_SEED": 2}),
]


@pytest.mark.parametrize("n,batch,reps,envs", CASES,
                         ids=[f"{e.get('AB_DIR', 'c2c')}-{n}x{b}-{e.get('AB_INPUT', 'sin')}"
                              f"{'-' + e['AB_VARIANT'] if 'AB_VARIANT' in e else ''}"
                              f"{'-inv' if e.get('AB_SGN') == '+1' else ''}" for n, b, _, e in CASES])
def test_cases_pass_and_report(tmp_path, n, batch, reps, envs):
    prefix = tmp_path / "d"
    r = run(n, batch, reps, AB_DUMP=prefix, **envs)
    assert r.returncode == 0, r.stdout + r.stderr
    rep = report.Report.parse(r.stdout)
    assert rep.verdict == "PASS" and rep.max_rel <= 1e-4
    kind = envs.get("AB_DIR", "c2c")
    is_long = kind == "c2c" and n > 4096
    assert rep.mode == ("staged" if is_long else "resident")
    # one-decimal printing: allow 0.1 us per launch span beside 1 %
    assert rep.problems(report.Tolerance(segments_us=0.25)) == []
    if is_long:
        assert rep.chain == ("fused2" if rep.launches == 2 else "single")
    else:
        assert rep.segments is None and rep.chain is None
    # the dumped output against the FP64 oracle
    x = np.fromfile(f"{prefix}.in.bin", np.float32)
    y = np.fromfile(f"{prefix}.out.bin", np.float32)
    sign = 1 if envs.get("AB_SGN") == "+1" else -1
    want = _oracle_ref(kind, x, n, sign)
    assert oracle.rel_l2_rows(y.reshape(want.shape), want) <= REL_L2


@pytest.mark.parametrize("n,batch", [(1024, 16), (8192, 2), (65536, 1)])
def test_input_sequence_reuses_the_plan_without_stale_outputs(n, batch):
    r = run(n, batch, 1, AB_INPUT_SEQ="sin,random-seeded,sin,impulse", AB_SEED=11)
    assert r.returncode == 0, r.stdout + r.stderr
    rep = report.Report.parse(r.stdout)
    assert [(s.verdict, s.stale) for s in rep.seq] == [("PASS", False)] * 4
    assert "stale-check on -> PASS" in r.stdout


@pytest.mark.parametrize("n,mode,host", [(4096, "async", "pinned"), (4096, "sync", "pageable"),
                                         (4096, "xfer", "pinned"), (65536, "async", "pinned"),
                                         (4096, "pipe", "pinned"), (16384, "pipe", "pageable"),
                                         (65536, "pipe", "pinned")])
def test_end_to_end_lines(n, mode, host):
    # pipe: bfly_host_pipelined in chunks of 3 transforms (a ragged last one),
    # its output checked against FP64 by the program itself
    r = run(n, 64 if n <= 4096 else 4, 3, AB_E2E=5, AB_E2E_MODE=mode, AB_E2E_HOST=host, AB_E2E_CHUNK=3)
    assert r.returncode == 0, r.stdout + r.stderr
    line = next(l for l in r.stdout.splitlines() if l.startswith("E2E n="))
    assert f"mode={mode} host={host}" in line
    e2e = float(line.split("e2e_us=")[1].split()[0])
    assert e2e > 0
    assert "E2E transfers: in=1 out=1 boundary=0 per_execution" in r.stdout
    if mode == "pipe":
        assert next(l for l in r.stdout.splitlines() if l.startswith("E2E pipe maxRel=")).endswith("PASS")
