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


@pytest.fixture(scope="module")
def codec_host(tmp_path_factory):
    """tests/native/codec_host.cpp: the kernels' per-block encoders and
    decoders (block_codec.h, block_decode.h) built for the host, with
    v_perm_b32 / v_lerp_u8 written from their documented semantics."""
    if not CLANG.exists():
        pytest.skip("clang++ of the ROCm toolchain not available")
    exe = tmp_path_factory.mktemp("codec") / "codec_host"
    subprocess.run([str(CLANG), "-std=c++17", "-O2", f"-I{ROOT / 'tests' / 'native' / 'host_shim'}",
                    f"-I{ROOT / 'goofy_amd' / 'csrc'}", str(ROOT / "tests" / "native" / "codec_host.cpp"),
                    "-o", str(exe)], check=True)
    return exe


@pytest.mark.parametrize("kind", ["uniform", "narrow", "ramps", "threshold", "alpha"])
def test_device_block_codec_on_host_matches_oracle(tmp_path, oracle, codec_host, kind):
    """All four encoders of the kernels against the oracle, block for block."""
    import numpy as np

    from blockgen import random_blocks
    n = 200000
    blocks = random_blocks(kind, 99, n)
    (tmp_path / "in.bin").write_bytes(blocks.tobytes())
    subprocess.run([str(codec_host), "encode", str(tmp_path / "in.bin"), str(tmp_path / "out.bin")], check=True,
                   timeout=120)
    out = np.frombuffer((tmp_path / "out.bin").read_bytes(), np.uint8)
    sizes = [8 * n, 8 * n, 16 * n, 16 * n]
    parts = np.split(out, np.cumsum(sizes)[:-1])
    for fmt, got in enumerate(parts):
        want = oracle.encode_blocks(fmt, blocks)
        bad = np.nonzero((got.reshape(n, -1) != want).any(1))[0]
        assert bad.size == 0, f"format {fmt}: {bad.size} blocks differ, first {bad[:5]}"


@pytest.mark.parametrize("fmt", [0, 1, 2, 3])
def test_device_block_decoders_on_host_match_oracle(tmp_path, oracle, harness, codec_host, fmt):
    """The kernels' decoders against the oracle's and the harness's on random
    blocks: every bit pattern, so for the ETC formats all five colour modes
    (individual, differential, and ETC2's T, H and planar, about 3% each);
    the second half of the ETC colour blocks is encoder output, half of it
    switched to individual mode."""
    import numpy as np

    from blockgen import random_blocks
    rng = np.random.default_rng(40 + fmt)
    n = 100000
    bb = 16 if fmt >= 2 else 8
    blocks = rng.integers(0, 256, (n, bb), dtype=np.uint8)
    if fmt in (1, 3):
        colour = oracle.encode_blocks(1, random_blocks("uniform", 41, n // 2))
        colour[::2, 3] &= 0xFD  # clear the diff bit: individual mode
        blocks[n // 2:, bb - 8:] = colour
    (tmp_path / "in.bin").write_bytes(blocks.tobytes())
    subprocess.run([str(codec_host), "decode", str(fmt), str(tmp_path / "in.bin"), str(tmp_path / "out.bin")],
                   check=True, timeout=120)
    got = np.frombuffer((tmp_path / "out.bin").read_bytes(), np.uint8).reshape(n, 64)
    img = oracle.decode_image(fmt, blocks.reshape(-1), 4 * n, 4).reshape(4, n, 16).transpose(1, 0, 2)
    assert np.array_equal(got, img.reshape(n, 64))
    ref = harness.decompress(fmt, blocks.reshape(-1), 4 * n, 4).reshape(4, n, 16).transpose(1, 0, 2)
    assert np.array_equal(got, ref.reshape(n, 64))


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
def test_stripe_division_magic_on_host(tmp_path):
    """internal.h's launch splitting: every stripe of max_stripe_brows(bw) rows
    indexes its blocks exactly, for powers of two, their neighbours and
    non-power-of-two widths near 2^19 and 2^20 (tests/native/stripe_magic.cpp)."""
    exe = tmp_path / "stripe_magic"
    subprocess.run(["g++", "-std=c++17", "-O2", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
                    f"-I{ROOT / 'include'}", f"-I{ROOT / 'goofy_amd' / 'csrc'}",
                    str(ROOT / "tests" / "native" / "stripe_magic.cpp"), "-o", str(exe)], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "stripe magic ok" in r.stdout


HIPCC = Path("/opt/rocm/bin/hipcc")


@pytest.mark.skipif(not HIPCC.exists(), reason="hipcc not available")
@pytest.mark.parametrize("source", ["kernels.hip", "codec_extra.hip", "probe.hip"])
def test_device_code_has_no_v_ashr_pk_u8_i32(tmp_path, source):
    """No kernel may contain v_ashr_pk_u8_i32.  The compiler makes it from an
    arithmetic shift followed by a clamp to [0, 255] and then relies on bits
    31:16 of its result being zero, which gfx950 does not guarantee: both the
    generator's normal-map branch and the ETC2 planar decoder produced wrong
    bytes on the GPU through it.  The device code is written to avoid the
    pattern; this pins that on every build."""
    asm = tmp_path / "k.s"
    subprocess.run([str(HIPCC), "-O3", "-std=c++17", "--offload-arch=gfx950", "-Xclang", "-target-feature",
                    "-Xclang", "-packed-fp32-ops", f"-I{ROOT / 'include'}", f"-I{ROOT / 'goofy_amd' / 'csrc'}",
                    "--cuda-device-only", "-S", str(ROOT / "goofy_amd" / "csrc" / source), "-o", str(asm)],
                   check=True, capture_output=True)
    text = asm.read_text()
    assert "v_lerp_u8" in text or source != "kernels.hip"  # the dump is the real kernels
    assert "v_ashr_pk_u8_i32" not in text
