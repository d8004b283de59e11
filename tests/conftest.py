"""Shared test setup: paths, the `gpu` marker, seeded inputs."""
import os
import shutil
import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
PKG = ROOT / "ascend-fft_amd"
for p in (str(ROOT), str(PKG)):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libbfly_hip.so)")


def gpu_available() -> bool:
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def dev():
    import torch
    if not torch.cuda.is_available():
        pytest.fail("GPU test collected without a GPU")
    return torch.device("cuda:0")


@pytest.fixture(scope="session")
def stub_cli(tmp_path_factory):
    """The bfly_check program built against the C ABI's CPU stub (ASan +
    UBSan): every guard runs, no transform is computed."""
    if not shutil.which("g++"):
        pytest.skip("needs g++")
    stub = PKG / "lib" / "libbfly_stub.so"
    assert stub.exists(), "build first: make -C ascend-fft_amd"
    exe = tmp_path_factory.mktemp("cli") / "bfly_check_stub"
    subprocess.run(
        ["g++", "-std=c++17", "-O1", "-Wall", "-Wextra", "-fsanitize=address,undefined",
         f"-I{ROOT / 'include'}", f"-I{PKG / 'csrc'}", f"-I{PKG / 'csrc' / 'check'}",
         str(PKG / "csrc" / "check" / "bfly_check.cpp"), str(PKG / "csrc" / "check" / "inputs.cpp"),
         str(PKG / "csrc" / "framework" / "reference.cpp"),
         f"-L{stub.parent}", "-lbfly_stub", f"-Wl,-rpath,{stub.parent}", "-o", str(exe)],
        check=True, capture_output=True, text=True)
    return exe
