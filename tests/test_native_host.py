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


CLANG = Path("/opt/rocm/llvm/bin/clang++")


@pytest.mark.skipif(not CLANG.exists(), reason="clang++ of the ROCm toolchain not available")
@pytest.mark.parametrize("kind", ["uniform", "narrow", "ramps", "threshold", "alpha"])
def test_device_block_codec_on_host_matches_oracle(tmp_path, oracle, kind):
    """goofy_amd/csrc/block_codec.h (the kernels' per-block arithmetic) built
    for the host, with v_perm_b32 / v_lerp_u8 written from their documented
    semantics: all four encoders against the oracle, block for block."""
    import numpy as np

    from blockgen import random_blocks
    exe = tmp_path / "codec_host"
    subprocess.run([str(CLANG), "-std=c++17", "-O2", f"-I{ROOT / 'tests' / 'native' / 'host_shim'}",
                    f"-I{ROOT / 'goofy_amd' / 'csrc'}", str(ROOT / "tests" / "native" / "codec_host.cpp"),
                    "-o", str(exe)], check=True)
    n = 200000
    blocks = random_blocks(kind, 99, n)
    (tmp_path / "in.bin").write_bytes(blocks.tobytes())
    subprocess.run([str(exe), str(tmp_path / "in.bin"), str(tmp_path / "out.bin")], check=True, timeout=120)
    out = np.frombuffer((tmp_path / "out.bin").read_bytes(), np.uint8)
    sizes = [8 * n, 8 * n, 16 * n, 16 * n]
    parts = np.split(out, np.cumsum(sizes)[:-1])
    for fmt, got in enumerate(parts):
        want = oracle.encode_blocks(fmt, blocks)
        bad = np.nonzero((got.reshape(n, -1) != want).any(1))[0]
        assert bad.size == 0, f"format {fmt}: {bad.size} blocks differ, first {bad[:5]}"
