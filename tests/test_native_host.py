"""Host-side native code under ThreadSanitizer (CPU only, no HIP)."""
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
def test_copy_pool_under_tsan(tmp_path):
    exe = tmp_path / "copy_pool_stress"
    subprocess.run(["g++", "-std=c++17", "-O1", "-g", "-fsanitize=thread", "-pthread",
                    f"-I{ROOT / 'goofy_amd' / 'csrc'}", str(ROOT / "tests" / "native" / "copy_pool_stress.cpp"),
                    "-o", str(exe)], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "WARNING: ThreadSanitizer" not in r.stderr, r.stderr
    assert "mismatches=0" in r.stdout


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc not available")
def test_oracle_under_asan_ubsan(tmp_path):
    exe = tmp_path / "oracle_sanitize"
    subprocess.run(["gcc", "-std=c11", "-O1", "-g", "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
                    f"-I{ROOT / 'oracle'}", str(ROOT / "tests" / "native" / "oracle_sanitize.c"),
                    str(ROOT / "oracle" / "goofy_oracle.c"), str(ROOT / "oracle" / "goofy_decode_oracle.c"),
                    "-lpthread", "-o", str(exe)], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "oracle sanitize ok" in r.stdout
