"""The render server's scheduler and admission controller (pure host C++,
axiomr_amd/csrc/serve_core.hpp) against a fake device, built and run under
AddressSanitizer + UBSan (g++) and ThreadSanitizer (clang++: its runtime
intercepts the pthread_cond_clockwait behind std::condition_variable's timed
waits, which GCC 11's does not)."""
import shutil
import subprocess
from pathlib import Path

import pytest

SRC = Path(__file__).resolve().parent / "native" / "test_serve_core.cpp"
CLANG = Path("/opt/rocm/llvm/bin/clang++")


def _build_and_run(tmp_path, compiler, flags, name):
    exe = tmp_path / name
    cmd = [str(compiler), "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer", "-pthread", *flags, "-o", str(exe),
           str(SRC)]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0 and "ALL OK" in r.stdout, r.stdout + r.stderr
    assert "WARNING" not in r.stderr and "ERROR" not in r.stderr, r.stderr


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not found")
def test_serve_core_asan_ubsan(tmp_path):
    _build_and_run(tmp_path, "g++", ["-fsanitize=address,undefined", "-fno-sanitize-recover=undefined"], "asan")


@pytest.mark.skipif(not CLANG.exists(), reason="clang++ with the TSan runtime not found")
def test_serve_core_tsan(tmp_path):
    _build_and_run(tmp_path, CLANG, ["-fsanitize=thread"], "tsan")
