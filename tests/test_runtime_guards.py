"""Compile the C++ Context / Plan runtime against the CPU stub of the C ABI
and run its host checks (tests/cpp/plan_host_checks.cpp): configuration and
guard return codes, cache invalidation after a failed table rebuild, launch
and copy failure codes, and the kernel every candidate lowers to.  The
reference runs its guard tests the same way on its fake AscendCL
(tests/test_runtime_guards.py of the reference)."""
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
PKG = ROOT / "ascend-fft_amd"


@pytest.mark.skipif(not shutil.which("g++"), reason="needs g++")
def test_plan_host_checks(tmp_path):
    stub = PKG / "lib" / "libbfly_stub.so"
    assert stub.exists(), "build first: make -C ascend-fft_amd"
    exe = tmp_path / "plan_host_checks"
    subprocess.run(
        ["g++", "-std=c++17", "-O1", "-Wall", "-Wextra", "-fsanitize=address,undefined",
         f"-I{ROOT / 'include'}", f"-I{PKG / 'csrc'}", f"-I{PKG / 'csrc' / 'framework'}",
         str(ROOT / "tests" / "cpp" / "plan_host_checks.cpp"),
         str(PKG / "csrc" / "framework" / "butterfly.cpp"),
         str(PKG / "csrc" / "framework" / "reference.cpp"),
         f"-L{stub.parent}", "-lbfly_stub", f"-Wl,-rpath,{stub.parent}", "-o", str(exe)],
        check=True, capture_output=True, text=True)
    r = subprocess.run([str(exe), str(PKG / "config"), str(tmp_path)], capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "plan host checks: 0 of" in r.stdout
    # the C++ reading of the fitted cost model equals the calibration's own
    # prediction (same key, same features, same coefficients)
    import json
    import math
    from bfly import calibrate
    model = json.loads(calibrate.MODEL_FILE.read_text())["kernels"]
    lines = [l.split() for l in r.stdout.splitlines() if l.startswith("eta ")]
    assert len(lines) == 8
    for _, kernel, n, b, eta in lines:
        n, b = int(n), int(b)
        c = model[calibrate.model_key(kernel, n)]["c"]
        want = sum(ci * fi for ci, fi in zip(c, calibrate.features(kernel, n, b)))
        assert math.isclose(float(eta), want, rel_tol=1e-6), (kernel, n, b, eta, want)
