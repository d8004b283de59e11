"""GPU: the drop-in C++ API (bfly::Context / bfly::Plan) end to end.

ascend-fft_amd/lib/bfly_plan_cases (tests/cpp/plan_gpu_cases.cpp, built by
the library's Makefile) selects plans through Context::select and runs
Plan::run / runInverse / runR2C / runC2R / runDevice / measure on the GPU with
host buffers, as a caller of the reference's API does.  Its own lines carry
the reference's metric (maxRelScaled, gate 1e-4) against the library's FP64
reference; the dumped outputs are gated here against the FP64 oracle
(relative L2 per row <= 1e-5).
"""
import re
import subprocess
from pathlib import Path

import numpy as np
import pytest

import oracle
from conftest import gpu_available

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not gpu_available(), reason="needs an MI355X")]

ROOT = Path(__file__).resolve().parents[1]
EXE = ROOT / "ascend-fft_amd" / "lib" / "bfly_plan_cases"
REL_L2 = 1e-5


@pytest.fixture(scope="module")
def run(tmp_path_factory):
    if not EXE.exists():
        pytest.fail("bfly_plan_cases is not built (make -C ascend-fft_amd)")
    out = tmp_path_factory.mktemp("plan")
    r = subprocess.run([str(EXE), str(ROOT / "ascend-fft_amd" / "config"), str(out)],
                       capture_output=True, text=True, timeout=240)
    print(r.stdout)
    return r, out


def test_every_call_returns_its_code_and_meets_the_reference_gate(run):
    r, _ = run
    assert r.returncode == 0, r.stdout + r.stderr
    assert "plan gpu cases: passed" in r.stdout
    cases = re.findall(r"case (\S+) n (\d+) batch (\d+) rc (-?\d+) maxrel (\S+)", r.stdout)
    assert len(cases) >= 20
    for name, n, b, rc, err in cases:
        if name.startswith("guard"):
            assert int(rc) == -8, (name, n)
            continue
        assert int(rc) == 0, (name, n, b)
        # the reference's gate; the inverse round trip: absolute error per
        # element of inputs in [-1, 1)
        assert float(err) <= (1e-4 if "round_trip" not in name else 2e-5), (name, n, b, err)
    m = re.search(r"case measure n 4096 batch 33 rc 0 us (\S+) maxrel (\S+) ok 1", r.stdout)
    assert m and float(m.group(1)) > 0 and float(m.group(2)) <= 1e-4


def _load(d, tag, rows):
    a = np.fromfile(d / f"{tag}_in.f32", np.float32)
    b = np.fromfile(d / f"{tag}_out.f32", np.float32)
    return a.reshape(rows, -1), b.reshape(rows, -1)


@pytest.mark.parametrize("n,b", [(64, 5), (1024, 64), (4096, 16), (8192, 3), (16384, 2), (65536, 2)])
def test_plan_run_c2c_against_fp64_oracle(run, n, b):
    _, d = run
    x, y = _load(d, f"c2c_n{n}_b{b}", b)
    assert oracle.rel_l2_rows(y, oracle.c2c_f64(x, n, -1)) <= REL_L2


@pytest.mark.parametrize("n", [128, 1024, 4096, 8192])
def test_plan_run_real_against_fp64_oracle(run, n):
    _, d = run
    x, h = _load(d, f"r2c_n{n}_b7", 7)
    assert oracle.rel_l2_rows(h, oracle.r2c_f64(x, n)) <= REL_L2
    if n <= 4096:
        h2, y = _load(d, f"c2r_n{n}_b7", 7)
        assert np.array_equal(h2, h)
        assert oracle.rel_l2_rows(y, oracle.c2r_f64(h2, n)) <= REL_L2
        assert np.max(np.abs(y - x)) < 1e-5


def test_select_measures_every_kernel_and_keeps_the_fastest(run):
    """Context::select at 4096 x 4096 measures each of the eight kernels the
    space lowers to there (split MFMA included), each once, and returns the
    fastest one that passed the gate."""
    r, _ = run
    log = [l[len("selectlog "):] for l in r.stdout.splitlines() if l.startswith("selectlog ")]
    measured = {}
    for l in log:
        m = re.match(r"\s+(\S+)\s+eta=\s*(\S+) us\s+measured=\s*(\S+) us\s+maxRel=(\S+)", l)
        if m:
            assert m.group(1) not in measured, l
            measured[m.group(1)] = float(m.group(3))
            assert float(m.group(4)) <= 1e-4
    assert set(measured) == {"plain", "stream", "plain_swz", "stream_swz", "stream_lt", "stream2",
                             "split_mfma", "split_mfma_stream"}, log
    assert len(set(measured.values())) >= 6                    # distinct measurements
    sel = re.search(r"case select n 4096 batch 4096 kernel (\S+) us (\S+)", r.stdout)
    assert sel and measured[sel.group(1)] == min(measured.values()), (sel and sel.group(0), measured)
