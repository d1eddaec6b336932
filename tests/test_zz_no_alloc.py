"""The allocation-free quality entry point (goofy_hip_quality_device_noalloc)
against goofy_hip_quality_device, through the C-ABI: every figure equal, on
two calls in a row, the second on a smaller image (the fixed device array
keeps the first call's sums and partials, which the second must not see)."""
import numpy as np
import pytest

from goofy_amd import workloads as W

pytestmark = pytest.mark.gpu


def _case(gpu, hip, fmt, w, h, pitch, seed):
    img = gpu.synth_host(2 if fmt >= W.DXT5 else 0, seed, 0, 20, w, h, pitch)
    src = hip.DeviceBuffer.from_array(img)
    blocks = hip.DeviceBuffer((w // 4) * (h // 4) * gpu.block_bytes(fmt))
    gpu.compress_device(fmt, blocks, src, w, h, pitch)
    hip.synchronize()
    return src, blocks


@pytest.mark.parametrize("fmt", [W.DXT1, W.ETC1, W.DXT5, W.ETC2])
def test_noalloc_quality_matches_pooled_call(gpu, fmt):
    from goofy_amd import hip
    # a large image (the full 1536-workgroup grid of partials), then a small one
    # (a few workgroups: partials beyond its grid still hold the first call's)
    for w, h, pitch, seed in ((2048, 1536, 2048 * 4 + 64, 11), (96, 40, 96 * 4, 12)):
        print(f"noalloc quality fmt={fmt} {w}x{h} pitch={pitch}", flush=True)
        src, blocks = _case(gpu, hip, fmt, w, h, pitch, seed)
        got = gpu.quality_device_noalloc(fmt, blocks, src, w, h, pitch)
        want = gpu.quality_device(fmt, blocks, src, w, h, pitch)
        assert got == want
        assert got["pixels"] == w * h and sum(got["sse_byte"]) > 0


def test_noalloc_quality_on_a_stream_and_errors(gpu):
    """On a caller's stream, twice in a row with the smaller image second;
    the checks of the pooled call (bad format, width) stay."""
    from goofy_amd import hip
    s = hip.Stream()
    a = _case(gpu, hip, W.ETC2, 1024, 512, 1024 * 4, 21)
    b = _case(gpu, hip, W.ETC2, 64, 8, 64 * 4, 22)
    for (src, blocks), (w, h) in ((a, (1024, 512)), (b, (64, 8))):
        print(f"noalloc quality on a stream {w}x{h}", flush=True)
        got = gpu.quality_device_noalloc(W.ETC2, blocks, src, w, h, w * 4, s)
        assert got == gpu.quality_device(W.ETC2, blocks, src, w, h, w * 4, s)
    src, blocks = a
    with pytest.raises(gpu.GoofyError) as e:
        gpu.quality_device_noalloc(9, blocks, src, 1024, 512)
    assert e.value.status == gpu.ERR_INVALID
    with pytest.raises(gpu.GoofyError) as e:
        gpu.quality_device_noalloc(W.ETC2, blocks, src, 1022, 512)
    assert e.value.status == gpu.ERR_WIDTH
    s.destroy()
