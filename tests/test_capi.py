"""The C-ABI library without a GPU: it loads, exports every declared symbol,
and rejects bad arguments with the reference's codes before touching a device."""
import ctypes
import re
import subprocess
from pathlib import Path

import numpy as np
import pytest

import goofy_amd

ROOT = Path(__file__).resolve().parents[1]
LIB = ROOT / "goofy_amd" / "libgoofy_hip.so"


def declared_functions():
    names = []
    for h in sorted((ROOT / "include").glob("*.h")):
        text = h.read_text()
        names += re.findall(r"GOOFY_HIP_API\s+[\w\s\*]+?\b(\w+)\s*\(", text)
    return names


def test_every_declared_symbol_is_exported():
    names = declared_functions()
    assert len(names) >= 15
    lib = ctypes.CDLL(str(LIB))
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_headers_compile_as_c_and_cpp(tmp_path):
    # include/goofy_hip.h is a C header (plain pointers and sizes); goofy_tc.h is
    # the reference's C++ API; both must also compile side by side in one TU
    c = tmp_path / "t.c"
    c.write_text('#include "goofy_hip.h"\nint main(void) { goofy_launch_config k = {1, 8, 0, 1, 0};'
                 ' goofy_quality q; (void)k; (void)q; return goofy_hip_device_count() < 0; }\n')
    subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-fsyntax-only", f"-I{ROOT / 'include'}", str(c)],
                   check=True)
    cpp = tmp_path / "t.cpp"
    cpp.write_text('#include "goofy_tc.h"\n#include "goofy_hip.h"\n'
                   'int main() { return goofy::compressDXT1(nullptr, nullptr, 0, 0, 0); }\n')
    subprocess.run(["g++", "-std=c++17", "-Wall", "-Werror", "-fsyntax-only", f"-I{ROOT / 'include'}", str(cpp)],
                   check=True)


def test_dropin_cpp_api_exported():
    out = subprocess.run(["nm", "-D", "-C", "--defined-only", str(LIB)], capture_output=True, text=True,
                         check=True).stdout
    for name in ("DXT1", "ETC1", "DXT5", "ETC2"):
        sig = f"goofy::compress{name}(unsigned char*, unsigned char const*, unsigned int, unsigned int, unsigned int)"
        assert sig in out, sig


def test_dropin_header_matches_reference_signature():
    text = (ROOT / "include" / "goofy_tc.h").read_text()
    for name in ("compressDXT1", "compressETC1"):
        assert re.search(rf"int {name}\(unsigned char\* result, const unsigned char\* input, unsigned int width,"
                         rf"\s+unsigned int height,\s+unsigned int stride\);", text), name


@pytest.mark.parametrize("fn", [goofy_amd.compressDXT1, goofy_amd.compressETC1])
def test_dropin_return_codes_without_device(fn):
    out = np.zeros(64, np.uint8)
    img = np.zeros(64 * 64 * 4, np.uint8)
    assert fn(out, img, 60, 64, 240) == -1   # width % 16 (GoofyTC/goofy_tc.h:1500)
    assert fn(out, img, 64, 62, 256) == -2   # height % 4 (:1505)
    assert fn(out, img, 60, 62, 240) == -1   # width first
    assert fn(out, img, 0, 64, 0) == 0       # nothing to do
    assert fn(out, img, 64, 0, 256) == 0


@pytest.mark.parametrize("fn", [goofy_amd.compressDXT5, goofy_amd.compressETC2])
def test_alpha_dropin_return_codes_without_device(fn):
    out = np.zeros(64, np.uint8)
    img = np.zeros(64 * 64 * 4, np.uint8)
    assert fn(out, img, 60, 64, 240) == -1
    assert fn(out, img, 64, 62, 256) == -2
    assert fn(out, img, 0, 64, 0) == 0


def test_format_block_bytes():
    f = goofy_amd._lib.goofy_format_block_bytes
    assert [f(k) for k in (0, 1, 2, 3, 4, -1)] == [8, 8, 16, 16, 0, 0]
    assert [goofy_amd.block_bytes(k) for k in (0, 1, 2, 3)] == [8, 8, 16, 16]


def test_dual_alpha_return_codes_without_device():
    out_a, out_b = np.zeros(64, np.uint8), np.zeros(64, np.uint8)
    img = np.zeros(64 * 64 * 4, np.uint8)
    f = goofy_amd._lib.goofy_hip_compress_dual_alpha
    u8 = goofy_amd._u8
    assert f(u8(out_a), u8(out_b), u8(img), 60, 64, 240) == -1
    assert f(u8(out_a), u8(out_b), u8(img), 64, 62, 256) == -2
    assert f(u8(out_a), u8(None), u8(img), 64, 64, 256) == goofy_amd.ERR_INVALID
    with pytest.raises(goofy_amd.GoofyError) as e:  # 16-byte blocks need 16-byte aligned outputs
        goofy_amd.compress_device_dual_alpha(16, 24, 16, 8, 4, 32)
    assert e.value.status == goofy_amd.ERR_INVALID


def test_dual_dropin_return_codes_without_device():
    out_a, out_b = np.zeros(64, np.uint8), np.zeros(64, np.uint8)
    img = np.zeros(64 * 64 * 4, np.uint8)
    f = goofy_amd._lib.goofy_hip_compress_dual
    u8 = goofy_amd._u8
    assert f(u8(out_a), u8(out_b), u8(img), 60, 64, 240) == -1
    assert f(u8(out_a), u8(out_b), u8(img), 64, 62, 256) == -2
    assert f(u8(out_a), u8(out_b), u8(img), 0, 64, 0) == 0
    assert f(u8(out_a), u8(None), u8(img), 64, 64, 256) == goofy_amd.ERR_INVALID


def test_compress_func_type_matches_reference_plugin():
    # CompressFunc_t (Src/main.cpp:644): int(uchar* dst, const uchar* src, uint w, uint h, uint stride)
    out = np.zeros(64, np.uint8)
    img = np.zeros(16 * 4 * 4, np.uint8)
    p = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))  # noqa: E731
    assert goofy_amd.compress_dxt1_fn(p(out), p(img), 12, 4, 48) == -1
    assert goofy_amd.compress_etc1_fn(p(out), p(img), 16, 3, 64) == -2


def test_device_api_validation_without_device():
    with pytest.raises(goofy_amd.GoofyError) as e:
        goofy_amd.compress_device(goofy_amd.DXT1, 8, 16, 6, 4, 32)
    assert e.value.status == goofy_amd.ERR_WIDTH
    with pytest.raises(goofy_amd.GoofyError) as e:
        goofy_amd.compress_device(goofy_amd.DXT1, 8, 16, 8, 6, 32)
    assert e.value.status == goofy_amd.ERR_HEIGHT
    for args in [(goofy_amd.DXT1, 8, 16, 8, 4, 24),      # pitch not a multiple of 16
                 (goofy_amd.DXT1, 8, 16, 8, 4, 16),      # pitch < 4*width
                 (goofy_amd.DXT1, 8, 20, 8, 4, 32),      # misaligned input
                 (goofy_amd.DXT1, 0, 16, 8, 4, 32),      # null output
                 (goofy_amd.DXT5, 8, 16, 8, 4, 32),      # 16-byte blocks need a 16-byte aligned output
                 (goofy_amd.ETC2, 24, 16, 8, 4, 32),
                 (4, 16, 16, 8, 4, 32),                  # unknown format
                 (7, 8, 16, 8, 4, 32)]:
        with pytest.raises(goofy_amd.GoofyError) as e:
            goofy_amd.compress_device(*args)
        assert e.value.status == goofy_amd.ERR_INVALID, args
    with pytest.raises(goofy_amd.GoofyError) as e:
        goofy_amd.compress_device(goofy_amd.DXT1, 8, 16, 4, 4, 16, config={"blocks_per_lane": 1, "wg_waves": 2,
                                                                            "waves_per_cu": 0, "nontemporal": 1,
                                                                            "tile_order": 2})
    assert e.value.status == goofy_amd.ERR_INVALID
    with pytest.raises(goofy_amd.GoofyError) as e:
        goofy_amd.compress_device(goofy_amd.DXT1, 8, 16, 4, 4, 16, config={"blocks_per_lane": 3, "wg_waves": 4,
                                                                            "waves_per_cu": 0, "nontemporal": 0})
    assert e.value.status == goofy_amd.ERR_INVALID


def test_floor_and_image_quality_validation_without_device():
    with pytest.raises(goofy_amd.GoofyError) as e:
        goofy_amd.rgb565_baseline_device(16, 16, 6, 4)
    assert e.value.status == goofy_amd.ERR_WIDTH
    with pytest.raises(goofy_amd.GoofyError) as e:
        goofy_amd.rgb565_baseline_device(16, 16, 8, 3)
    assert e.value.status == goofy_amd.ERR_HEIGHT
    for args in [(16, 16, 8, 4, 24, 32),   # dst pitch not a multiple of 16
                 (16, 16, 8, 4, 32, 16),   # src pitch < 4*width
                 (16, 24, 8, 4, 32, 32)]:  # misaligned source
        with pytest.raises(goofy_amd.GoofyError) as e:
            goofy_amd.rgb565_baseline_device(*args)
        assert e.value.status == goofy_amd.ERR_INVALID, args
    with pytest.raises(goofy_amd.GoofyError) as e:
        goofy_amd.quality_images_device(16, 0, 8, 4)
    assert e.value.status == goofy_amd.ERR_INVALID
    with pytest.raises(goofy_amd.GoofyError) as e:
        goofy_amd.quality_images_device(16, 16, 10, 4)
    assert e.value.status == goofy_amd.ERR_WIDTH


def test_dual_validation_without_device():
    for args, status in [((8, 8, 16, 6, 4, 32), goofy_amd.ERR_WIDTH),
                         ((8, 8, 16, 8, 6, 32), goofy_amd.ERR_HEIGHT),
                         ((8, 8, 16, 8, 4, 24), goofy_amd.ERR_INVALID),     # pitch % 16
                         ((8, 12, 16, 8, 4, 32), goofy_amd.ERR_INVALID),    # misaligned ETC1S output
                         ((8, 0, 16, 8, 4, 32), goofy_amd.ERR_INVALID),     # null output
                         ((8, 8, 16, 4194304, 4, 4194304 * 4), goofy_amd.ERR_TOO_LARGE)]:
        with pytest.raises(goofy_amd.GoofyError) as e:
            goofy_amd.compress_device_dual(*args)
        assert e.value.status == status, args


def test_too_wide_images_refused_without_device():
    # 4,194,304 pixels per row (2^20 blocks) is beyond one launch's index range
    # even for a single block row; narrower images of any height are striped
    with pytest.raises(goofy_amd.GoofyError) as e:
        goofy_amd.compress_device(goofy_amd.DXT1, 8, 16, 4194304, 4, 4194304 * 4)
    assert e.value.status == goofy_amd.ERR_TOO_LARGE
    with pytest.raises(goofy_amd.GoofyError) as e:
        goofy_amd.Batch([dict(src=16, dst=8, width=4194304, height=4, pitch=4194304 * 4, format=0)])
    assert e.value.status == goofy_amd.ERR_TOO_LARGE


def test_probe_validation_without_device():
    with pytest.raises(goofy_amd.GoofyError) as e:
        goofy_amd.traffic_probe(8, 16, 16, 4, 64, wg_waves=3)
    assert e.value.status == goofy_amd.ERR_INVALID and "wg_waves" in goofy_amd.last_error()
    with pytest.raises(goofy_amd.GoofyError) as e:
        goofy_amd.traffic_probe(8, 16, 18, 4, 80)
    assert e.value.status == goofy_amd.ERR_WIDTH


def test_batch_validation_without_device():
    bad = [dict(src=16, dst=8, width=6, height=4, pitch=32, format=0)]
    with pytest.raises(goofy_amd.GoofyError) as e:
        goofy_amd.Batch(bad)
    assert e.value.status == goofy_amd.ERR_WIDTH
    bad = [dict(src=16, dst=8, width=8, height=4, pitch=32, format=0),
           dict(src=16, dst=8, width=8, height=4, pitch=40, format=0)]
    with pytest.raises(goofy_amd.GoofyError) as e:
        goofy_amd.Batch(bad)
    assert e.value.status == goofy_amd.ERR_INVALID and "texture 1" in goofy_amd.last_error()
    bad = [dict(src=16, dst=16, width=8, height=4, pitch=32, format=goofy_amd.DXT5),
           dict(src=16, dst=8, width=8, height=4, pitch=32, format=goofy_amd.ETC2)]  # 16-byte blocks, 8-aligned
    with pytest.raises(goofy_amd.GoofyError) as e:
        goofy_amd.Batch(bad)
    assert e.value.status == goofy_amd.ERR_INVALID and "texture 1" in goofy_amd.last_error()


def test_no_device_is_reported():
    if goofy_amd.device_count() > 0:
        pytest.skip("a GPU is visible")
    with pytest.raises(goofy_amd.GoofyError) as e:
        goofy_amd.compress(goofy_amd.DXT1, np.zeros(16 * 4 * 4, np.uint8), 16, 4)
    assert e.value.status == goofy_amd.ERR_NO_DEVICE


def test_synth_host_deterministic_and_threaded():
    a = goofy_amd.synth_host(0, 77, 3, 20, 100, 52, threads=1)
    b = goofy_amd.synth_host(0, 77, 3, 20, 100, 52, threads=7)
    assert np.array_equal(a, b)
    c = goofy_amd.synth_host(1, 77, 3, 0, 130, 70, pitch=130 * 4 + 24, fill=0xEE)
    rows = c.reshape(70, 130 * 4 + 24)
    assert (rows[:, 130 * 4:] == 0xEE).all()          # padding untouched
    assert (rows[:, 3:130 * 4:4] == 255).all()          # alpha = 255


def test_bench_refuses_more_than_one_gpu():
    # bench.py measures one GPU; --gpus N > 1 (or a multi-rank launch) exits with a message
    import sys
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 2 and "one GPU" in r.stderr and r.stdout == ""
