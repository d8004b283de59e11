"""Compile the C++ Context/Plan runtime against the CPU stub of the C ABI and
run its guard tests (reference tests/test_runtime_guards.py does the same with
its fake AscendCL)."""
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
PKG = ROOT / "ascend-fft_amd"


@pytest.mark.skipif(not shutil.which("g++"), reason="needs g++")
def test_host_guards_and_failed_prepare(tmp_path):
    stub = PKG / "lib" / "libbfly_stub.so"
    assert stub.exists(), "build first: make -C ascend-fft_amd"
    exe = tmp_path / "runtime_guards"
    subprocess.run(
        ["g++", "-std=c++17", "-O1", "-Wall", "-Wextra", "-fsanitize=address,undefined",
         f"-I{ROOT / 'include'}", f"-I{PKG / 'csrc'}", f"-I{PKG / 'csrc' / 'framework'}",
         str(ROOT / "tests" / "cpp" / "test_runtime_guards.cpp"),
         str(PKG / "csrc" / "framework" / "butterfly.cpp"),
         str(PKG / "csrc" / "framework" / "reference.cpp"),
         f"-L{stub.parent}", "-lbfly_stub", f"-Wl,-rpath,{stub.parent}", "-o", str(exe)],
        check=True, capture_output=True, text=True)
    r = subprocess.run([str(exe), str(PKG / "config"), str(tmp_path)], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "checks passed" in r.stdout
