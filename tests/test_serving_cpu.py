"""The serving scheduler (goofy_amd/csrc/serve.cpp) on the CPU, on a fake
device (tests/native/serve_fake_backend.cpp: streams as worker threads, events,
failure injection), under ThreadSanitizer and under AddressSanitizer with
UndefinedBehaviorSanitizer.  The cases (tests/native/serve_fake_test.cpp):
concurrent clients with and without batching, budget and queue-limit
backpressure, cancellation, deadlines, wait_idle while outputs hold the
arena, a failing launch call and a faulting kernel, allocation failure at
creation, destroy with work in flight, and the Poisson load generator with
every result checked.  No GPU."""
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
NATIVE = ROOT / "tests" / "native"
SOURCES = [str(ROOT / "goofy_amd" / "csrc" / "serve.cpp"), str(NATIVE / "serve_fake_backend.cpp"),
           str(NATIVE / "serve_fake_test.cpp")]
INCLUDES = [f"-I{NATIVE / 'fake_hip'}", f"-I{ROOT / 'include'}", f"-I{ROOT / 'goofy_amd' / 'csrc'}"]
# gcc 11's ThreadSanitizer does not intercept pthread_cond_clockwait (which
# std::condition_variable::wait_for uses) and reports false races; the ROCm
# clang's runtime does.
CLANG = Path("/opt/rocm/llvm/bin/clang++")


def _build_and_run(tmp_path, compiler, flags, name):
    exe = tmp_path / name
    subprocess.run([compiler, "-std=c++17", "-O1", "-g", "-pthread", *flags, *INCLUDES, *SOURCES, "-o", str(exe)],
                   check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "serve fake ok" in r.stdout, r.stdout
    return r


@pytest.mark.skipif(not CLANG.exists(), reason="clang++ of the ROCm toolchain not available")
def test_scheduler_under_tsan(tmp_path):
    r = _build_and_run(tmp_path, str(CLANG), ["-fsanitize=thread"], "serve_tsan")
    assert "WARNING: ThreadSanitizer" not in r.stderr, r.stderr


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
def test_scheduler_under_asan_ubsan(tmp_path):
    r = _build_and_run(tmp_path, "g++", ["-fsanitize=address,undefined", "-fno-sanitize-recover=all"], "serve_asan")
    assert "ERROR: AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr, r.stderr
