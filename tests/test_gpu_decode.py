"""GPU BC1/ETC1 (and BC3 / ETC2 RGBA8) decoders and the round-trip quality
kernel, through the C-ABI.

Decoders: bit-exact against the decoder oracle, which is itself pinned to the
reference harness's decoders (tests/test_oracle.py).  Quality: integer
squared errors exact; the luma error is a double-precision sum in a different
order than the reference's serial loop, so it is compared with relative
tolerance 1e-9 (and the PSNR figures to 1e-7 dB).
"""
import math

import numpy as np
import pytest

from goofy_amd import workloads as W

from conftest import REAL_IMAGES

pytestmark = pytest.mark.gpu


def etc_modes(blocks):
    """Colour mode of each 8-byte ETC block: 0 individual, 1 differential,
    2 T, 3 H, 4 planar (ETC2's modes, signalled by a differential red, green
    or blue sum outside 0..31, checked in that order)."""
    b = blocks.astype(np.int32)
    diff = (b[:, 3] & 2) != 0
    over = [((b[:, c] >> 3) + ((b[:, c] & 7) - 8 * ((b[:, c] & 4) != 0)) < 0) |
            ((b[:, c] >> 3) + ((b[:, c] & 7) - 8 * ((b[:, c] & 4) != 0)) > 31) for c in range(3)]
    return np.select([~diff, over[0], over[1], over[2]], [0, 2, 3, 4], 1)


def random_etc1(rng, n):
    """Random 8-byte ETC colour blocks, every mode included.  Random bytes
    give about 3% each of T, H and planar blocks; 30% more are forced into
    one of the three by setting the differential bit, keeping the channels
    before the chosen one in range (base 4..27) and making the chosen one
    overflow (base 31 with delta +1..+3, or base 0 with delta -3..-1)."""
    b = rng.integers(0, 256, (n, 8), dtype=np.uint8)
    forced = np.nonzero(rng.random(n) < 0.3)[0]
    ch = rng.integers(0, 3, forced.size)
    low = rng.random(forced.size) < 0.5
    d3 = rng.integers(1, 4, forced.size)
    b[forced, 3] |= 2
    for j, i in enumerate(forced):
        for c in range(ch[j]):
            b[i, c] = (int(b[i, c]) & 7) | ((4 + int(b[i, c]) % 24) << 3)
        b[i, ch[j]] = (8 - d3[j]) if low[j] else ((31 << 3) | d3[j])
    modes = etc_modes(b)
    for m in (2, 3, 4):
        assert (modes == m).mean() > 0.08, (m, (modes == m).mean())
    return b


@pytest.mark.parametrize("fmt", [W.DXT1, W.ETC1])
def test_decode_matches_oracle(gpu, oracle, harness, fmt):
    """Random blocks: every BC1 mode, and every ETC colour mode (individual,
    differential, T, H, planar), against the oracle's and the harness's
    decoders."""
    from goofy_amd import hip
    rng = np.random.default_rng(7 + fmt)
    bw, bh = 100, 37
    n = bw * bh
    blocks = rng.integers(0, 256, (n, 8), dtype=np.uint8) if fmt == W.DXT1 else random_etc1(rng, n)
    w, h = 4 * bw, 4 * bh
    pitch = w * 4 + 48
    d_blocks = hip.DeviceBuffer.from_array(blocks.reshape(-1))
    d_img = hip.DeviceBuffer(pitch * h, fill=0x3C)
    print(f"decode fmt={fmt} {w}x{h} pitch={pitch}", flush=True)
    gpu.decode_device(fmt, d_img, d_blocks, w, h, pitch)
    hip.synchronize()
    got = d_img.download().reshape(h, pitch)
    want = oracle.decode_image(fmt, blocks.reshape(-1), w, h).reshape(h, w * 4)
    assert np.array_equal(got[:, : w * 4], want)
    assert np.array_equal(want.reshape(-1), harness.decompress(fmt, blocks.reshape(-1), w, h))
    assert (got[:, w * 4:] == 0x3C).all()  # padding untouched


@pytest.mark.parametrize("fmt", [W.DXT5, W.ETC2])
def test_decode_alpha_formats_match_oracle_and_harness(gpu, oracle, harness, fmt):
    """Random alpha blocks (every BC3 mode, every EAC table and multiplier,
    0 included) with random colour halves (every ETC2 colour mode), against
    the oracle's and the harness's decoders."""
    from goofy_amd import hip
    rng = np.random.default_rng(17 + fmt)
    bw, bh = 96, 40
    n = bw * bh
    alpha = rng.integers(0, 256, (n, 8), dtype=np.uint8)
    colour = rng.integers(0, 256, (n, 8), dtype=np.uint8) if fmt == W.DXT5 else random_etc1(rng, n)
    blocks = np.concatenate([alpha, colour], axis=1)
    w, h = 4 * bw, 4 * bh
    pitch = w * 4 + 32
    d_blocks = hip.DeviceBuffer.from_array(blocks.reshape(-1))
    d_img = hip.DeviceBuffer(pitch * h, fill=0x3C)
    print(f"decode fmt={fmt} {w}x{h} pitch={pitch}", flush=True)
    gpu.decode_device(fmt, d_img, d_blocks, w, h, pitch)
    hip.synchronize()
    got = d_img.download().reshape(h, pitch)
    want = oracle.decode_image(fmt, blocks.reshape(-1), w, h).reshape(h, w * 4)
    assert np.array_equal(got[:, : w * 4], want)
    assert np.array_equal(want.reshape(-1), harness.decompress(fmt, blocks.reshape(-1), w, h))
    assert (got[:, w * 4:] == 0x3C).all()


def reference_quality(src, dec):
    """getMsePsnr (Src/main.cpp:403-469) restated in numpy float64."""
    a = src.reshape(-1, 4).astype(np.float64)
    b = dec.reshape(-1, 4).astype(np.float64)
    n = a.shape[0]
    se = ((a - b) ** 2).sum(0)
    y1 = 0.21 * a[:, 2] + 0.72 * a[:, 1] + 0.07 * a[:, 0]
    y2 = 0.21 * b[:, 2] + 0.72 * b[:, 1] + 0.07 * b[:, 0]
    mse_y = ((y2 - y1) ** 2).sum() / n
    psnr = lambda m, p=255.0: float("inf") if m < 1e-7 else 10 * math.log10(p * p / m)  # noqa: E731
    return {"sse_byte": [int(x) for x in se], "mse_y": mse_y, "psnr_y": psnr(mse_y),
            "psnr_rgb": psnr(se[:3].sum() / n, 768.0), "psnr_r": psnr(se[2] / n)}


@pytest.mark.parametrize("fmt", [W.DXT1, W.ETC1])
def test_quality_matches_harness_definition(gpu, oracle, fmt):
    from goofy_amd import hip
    im = W.IMAGE_512
    img = gpu.synth_host(im.kind, im.seed, im.texture_id, im.noise, im.width, im.height, im.pitch)
    src = hip.DeviceBuffer.from_array(img)
    blocks = hip.DeviceBuffer(im.blocks * 8)
    gpu.compress_device(fmt, blocks, src, im.width, im.height, im.pitch)
    q = gpu.quality_device(fmt, blocks, src, im.width, im.height, im.pitch)
    enc = blocks.download()
    dec = oracle.decode_image(fmt, enc, im.width, im.height)
    ref = reference_quality(img, dec)
    print(f"quality fmt={fmt}: psnr_rgb={q['psnr_rgb']:.3f} psnr_y={q['psnr_y']:.3f}", flush=True)
    assert q["pixels"] == im.pixels
    assert q["sse_byte"] == ref["sse_byte"]
    assert q["mse_y"] == pytest.approx(ref["mse_y"], rel=1e-9)
    assert q["psnr_y"] == pytest.approx(ref["psnr_y"], abs=1e-7)
    assert q["psnr_rgb"] == pytest.approx(ref["psnr_rgb"], abs=1e-7)
    assert q["psnr_r"] == pytest.approx(ref["psnr_r"], abs=1e-7)
    assert q["mse_a"] == 0.0 and q["psnr_a"] > 1e300  # alpha is 255 in and out


@pytest.mark.parametrize("w,h,pitch", [(200, 84, 832), (1028, 36, 4160), (272, 8, (1 << 29) - 16), (272, 8, 1 << 29)])
def test_decode_and_quality_address_paths(gpu, oracle, w, h, pitch):
    """The decode and round-trip kernels on both sides of their addressing
    limits: 50 blocks a row (under a wave's 64: per-lane 64-bit addresses),
    257 (one buffer resource per wave, runs straddling two block rows), and
    68 with pitches just under and at 2^29 bytes, where the wave-uniform base
    gives way to per-lane addresses (about 3.7 GB each for source and
    decoded image).  Decoded rows against the oracle's decoder, sums against
    getMsePsnr restated."""
    from goofy_amd import hip
    img = gpu.synth_host(W.WAVES_ALPHA, 5 + w, 0, 20, w, h, w * 4)
    src = hip.DeviceBuffer(pitch * (h - 1) + w * 4)
    gpu.synth_device(W.WAVES_ALPHA, 5 + w, 0, 20, src, w, h, pitch)
    dst = hip.DeviceBuffer(pitch * (h - 1) + w * 4)
    try:
        for fmt in (W.DXT1, W.ETC2):
            n = (w // 4) * (h // 4) * gpu.block_bytes(fmt)
            blocks = hip.DeviceBuffer(n)
            print(f"address paths {w}x{h} pitch={pitch} fmt={fmt}", flush=True)
            gpu.compress_device(fmt, blocks, src, w, h, pitch)
            enc = blocks.download()
            assert np.array_equal(enc, oracle.compress_blocks(fmt, img, w, h, w * 4))
            gpu.decode_device(fmt, dst, blocks, w, h, pitch)
            hip.synchronize()
            want = oracle.decode_image(fmt, enc, w, h).reshape(h, w * 4)
            got = np.stack([dst.download(w * 4, y * pitch) for y in range(h)])
            assert np.array_equal(got, want)
            q = gpu.quality_device(fmt, blocks, src, w, h, pitch)
            assert q["sse_byte"] == reference_quality(img, want)["sse_byte"]
            blocks.free()
    finally:
        src.free()
        dst.free()


def integer_sums(src, dec):
    """The round-trip sums in integers: squared error of bytes 0..3, and
    10^4 x the harness's luma error sum (its weights 0.21 / 0.72 / 0.07 on
    bytes 2 / 1 / 0, as 21 / 72 / 7)."""
    a = src.reshape(-1, 4)
    b = dec.reshape(-1, 4)
    d = [a[:, c].astype(np.int32) - b[:, c].astype(np.int32) for c in range(4)]
    sse = [int((x.astype(np.int64) ** 2).sum()) for x in d]
    y = 21 * d[2] + 72 * d[1] + 7 * d[0]
    return sse, int((y.astype(np.int64) ** 2).sum())


@pytest.mark.parametrize("fmt", [W.DXT1, W.ETC1, W.DXT5, W.ETC2])
@pytest.mark.parametrize("w,h", [(2600, 2600), (3588, 3512)])
def test_quality_waves_of_several_tiles(gpu, oracle, fmt, w, h):
    """Images large enough that each wave of the round-trip kernel compares
    two or three tiles of 64 blocks (its grid is at most 1536 workgroups of
    four waves: a stride of 6144 tiles), with a partial last tile.  In 2600^2
    (422,500 blocks, 6602 tiles) that tile is a wave's second; in 3588x3512
    (787,566 blocks, 12,306 tiles) a wave's third.  BC1 / BC3 run these waves
    with two tiles in registers.  Sums against the oracle's decode of the same
    blocks, in integers."""
    from goofy_amd import hip
    kind = W.WAVES_ALPHA if fmt in (W.DXT5, W.ETC2) else W.WAVES
    img = gpu.synth_host(kind, 31 + fmt, 0, 16, w, h, w * 4)
    src = hip.DeviceBuffer.from_array(img)
    blocks = hip.DeviceBuffer((w // 4) * (h // 4) * gpu.block_bytes(fmt))
    try:
        print(f"quality waves of several tiles {w}x{h} fmt={fmt}", flush=True)
        gpu.compress_device(fmt, blocks, src, w, h, w * 4)
        q = gpu.quality_device(fmt, blocks, src, w, h, w * 4)
        dec = oracle.decode_image(fmt, blocks.download(), w, h)
    finally:
        src.free()
        blocks.free()
    sse, y100 = integer_sums(img, dec)
    assert q["pixels"] == w * h
    assert q["sse_byte"] == sse
    assert q["sse_luma"] == pytest.approx(y100 / 1e4, rel=1e-12)


@pytest.mark.parametrize("fmt", [W.DXT5, W.ETC2])
def test_quality_alpha_formats_match_harness(gpu, harness, fmt):
    """Round trip of the alpha image through the 16-byte formats: the GPU
    quality call against getMsePsnr of the harness over the harness's own
    decode of the same blocks."""
    from goofy_amd import hip
    im = W.ALPHA_512
    img = gpu.synth_host(im.kind, im.seed, im.texture_id, im.noise, im.width, im.height, im.pitch)
    src = hip.DeviceBuffer.from_array(img)
    blocks = hip.DeviceBuffer(im.blocks * 16)
    gpu.compress_device(fmt, blocks, src, im.width, im.height, im.pitch)
    q = gpu.quality_device(fmt, blocks, src, im.width, im.height, im.pitch)
    dec = harness.decompress(fmt, blocks.download(), im.width, im.height)
    ref = harness.mse_psnr(img, dec, im.width, im.height)
    ours = reference_quality(img, dec)
    print(f"quality fmt={fmt}: psnr_a={q['psnr_a']:.3f} psnr_rgb={q['psnr_rgb']:.3f}", flush=True)
    assert q["sse_byte"] == ours["sse_byte"] and q["sse_byte"][3] > 0
    for k in ("mse_r", "mse_g", "mse_b", "mse_a", "mse_rgb", "psnr_a", "psnr_min", "psnr_rgb"):
        assert q[k] == pytest.approx(ref[k], rel=1e-12), k
    assert q["mse_y"] == pytest.approx(ref["mse_y"], rel=1e-9)
    assert q["psnr_a"] > 40.0  # nearest-level alpha: error at most a few units per pixel


@pytest.mark.parametrize("fmt", [W.ETC1, W.ETC2])
def test_quality_of_foreign_etc_blocks(gpu, harness, fmt):
    """The round-trip sums over ETC blocks the encoder never writes (random
    colour halves in all five modes, random EAC alpha): the GPU quality call
    against getMsePsnr of the harness over the harness's own decode."""
    from goofy_amd import hip
    rng = np.random.default_rng(70 + fmt)
    bw, bh = 64, 48
    n = bw * bh
    colour = random_etc1(rng, n)
    blocks = colour if fmt == W.ETC1 else np.concatenate([rng.integers(0, 256, (n, 8), dtype=np.uint8), colour], 1)
    w, h = 4 * bw, 4 * bh
    img = rng.integers(0, 256, w * h * 4, dtype=np.uint8)
    print(f"quality of foreign ETC blocks fmt={fmt} {w}x{h}", flush=True)
    q = gpu.quality_device(fmt, hip.DeviceBuffer.from_array(blocks.reshape(-1)), hip.DeviceBuffer.from_array(img),
                           w, h, w * 4)
    dec = harness.decompress(fmt, blocks.reshape(-1), w, h)
    ref = harness.mse_psnr(img, dec, w, h)
    assert q["sse_byte"] == reference_quality(img, dec)["sse_byte"]
    for k in ("mse_r", "mse_g", "mse_b", "mse_a", "mse_rgb", "psnr_min", "psnr_rgb"):
        assert q[k] == pytest.approx(ref[k], rel=1e-12), k
    assert q["mse_y"] == pytest.approx(ref["mse_y"], rel=1e-9)


def test_quality_of_black_image_closed_form(gpu):
    from goofy_amd import hip
    # black 16x16: c0 = 0x0020 (the forced green LSB), c1 = 0, every index 2,
    # so every pixel decodes to (0, (2*4+0)/3, 0) = (0, 2, 0): squared error 4
    # on byte 1 only; alpha is exact
    img = np.tile(np.array([0, 0, 0, 255], np.uint8), 16 * 16)
    src = hip.DeviceBuffer.from_array(img)
    blocks = hip.DeviceBuffer(16 * 8)
    gpu.compress_device(W.DXT1, blocks, src, 16, 16, 64)
    q = gpu.quality_device(W.DXT1, blocks, src, 16, 16, 64)
    assert q["sse_byte"] == [0, 4 * 256, 0, 0]
    assert q["psnr_g"] == pytest.approx(10 * math.log10(255.0 ** 2 / 4.0), abs=1e-12)
    assert q["psnr_min"] == q["psnr_g"] and q["psnr_a"] > 1e300 and q["psnr_r"] > 1e300


# ---- the harness's RGB565 quality floor and image-to-image quality ------------------------------

def test_rgb565_floor_vs_oracle_with_pitches(gpu, oracle):
    from goofy_amd import hip
    rng = np.random.default_rng(11)
    w, h = 200, 34
    sp, dp = w * 4 + 48, w * 4 + 32
    src = rng.integers(0, 256, sp * h, dtype=np.uint8)
    d_src = hip.DeviceBuffer.from_array(src)
    d_dst = hip.DeviceBuffer(dp * h, fill=0x5A)
    print(f"rgb565 floor {w}x{h} pitches {sp}/{dp}", flush=True)
    gpu.rgb565_baseline_device(d_dst, d_src, w, h, dst_pitch=dp, src_pitch=sp)
    hip.synchronize()
    got = d_dst.download().reshape(h, dp)
    want = oracle.rgb565_baseline(src, w, h, pitch=sp).reshape(h, w * 4)
    assert np.array_equal(got[:, : w * 4], want)
    assert (got[:, w * 4:] == 0x5A).all()


def _check_quality(q, ref, n):
    # squared errors are integers and the MSEs come from them exactly as in
    # getMsePsnr; only the luma sum's order differs (relative 1e-9)
    for k in ("mse_r", "mse_g", "mse_b", "mse_a", "mse_max", "mse_rgb",
              "psnr_r", "psnr_g", "psnr_b", "psnr_a", "psnr_min", "psnr_rgb"):
        assert q[k] == ref[k], k
    assert q["pixels"] == n
    assert q["mse_y"] == pytest.approx(ref["mse_y"], rel=1e-9)
    assert q["psnr_y"] == pytest.approx(ref["psnr_y"], abs=1e-7)


@pytest.mark.parametrize("name", REAL_IMAGES)
def test_floor_and_quality_real_images_match_harness(gpu, real_images, name):
    """GPU floor, its PSNR, and the encoder's round-trip PSNR on the
    reference's test images against the harness's own figures (fixtures)."""
    import hashlib

    from conftest import GOLDEN
    from goofy_amd import hip
    from goofy_amd import io as gio
    g = real_images[name]
    img, w, h = gio.load_png_rgba8(GOLDEN / "images" / name)
    assert hashlib.sha256(img.tobytes()).hexdigest() == g["input_sha256"]
    print(f"floor+quality {name} {w}x{h}", flush=True)
    d_src = hip.DeviceBuffer.from_array(img)
    d_base = hip.DeviceBuffer(img.size)
    gpu.rgb565_baseline_device(d_base, d_src, w, h)
    hip.synchronize()
    assert hashlib.sha256(d_base.download().tobytes()).hexdigest() == g["baseline_sha256"]
    _check_quality(gpu.quality_images_device(d_src, d_base, w, h), g["baseline_quality"], w * h)
    blocks = hip.DeviceBuffer((w // 4) * (h // 4) * 8)
    for fmt, key in ((W.DXT1, "dxt1"), (W.ETC1, "etc1")):
        gpu.compress_device(fmt, blocks, d_src, w, h, w * 4)
        _check_quality(gpu.quality_device(fmt, blocks, d_src, w, h), g[f"{key}_quality"], w * h)


def test_quality_sums_past_one_launch_closed_form(gpu):
    """819,200,000 blocks (4096 x 3,200,000 pixels, a 52.4 GB source): more than
    the 805,306,368 blocks that one round-trip launch may cover with 32-bit lane
    sums (2048 blocks per lane), so the call runs as two stripes, the second
    adding onto the first.  A white source against all-zero DXT1 blocks (c0 =
    c1 = 0: 3-colour mode, index 0, opaque black) errs by 255 in bytes 0..2 of
    every pixel.  That is the largest per-block error, so the lanes' sums sit
    at the top of their range; in one unstriped launch (2084 blocks per lane)
    the signed cross-product sums would wrap.  Closed form: 65025 x pixels on
    bytes 0..2, 0 on alpha, and 65025 x pixels of luma error
    ((21 + 72 + 7) x 255 = 25500 per pixel, squared, / 10^4)."""
    from goofy_amd import hip
    w, h = 4096, 3_200_000
    pixels = w * h
    print(f"quality past one launch: {w}x{h}, {pixels // 16} blocks", flush=True)
    src = hip.DeviceBuffer(w * 4 * h, fill=0xFF)
    blocks = hip.DeviceBuffer((w // 4) * (h // 4) * 8, fill=0)
    try:
        q = gpu.quality_device(W.DXT1, blocks, src, w, h, w * 4)
    finally:
        src.free()
        blocks.free()
    assert q["pixels"] == pixels
    assert q["sse_byte"] == [65025 * pixels] * 3 + [0]
    assert q["sse_luma"] == 65025.0 * pixels
    assert q["psnr_a"] > 1e300 and q["psnr_rgb"] == pytest.approx(10 * math.log10(768.0 ** 2 / (3 * 65025.0)),
                                                                  abs=1e-9)


@pytest.mark.parametrize("fmt", [W.ETC1, W.ETC2])
def test_quality_sums_past_one_launch_etc_closed_form(gpu, oracle, fmt):
    """The same 819,200,000 blocks for the ETC round trips, which add two
    pixels per v_dot2_i32_i16 into signed 32-bit lane sums.  All-zero ETC1
    blocks decode to (2, 2, 2, 255) and ETC2 RGBA8 ones to (2, 2, 2, 0) (the
    oracle's decoder, asserted here), so a white source errs by 253 in bytes
    0..2 and, for ETC2 RGBA8, by 255 in alpha: a lane's alpha sum reaches
    2048 x 16 x 65025 = 2,130,739,200, just under 2^31."""
    from goofy_amd import hip
    bb = gpu.block_bytes(fmt)
    px = oracle.decode_image(fmt, np.zeros(bb, np.uint8), 4, 4).reshape(16, 4)
    assert (px == px[0]).all()
    want_px = [2, 2, 2, 255] if fmt == W.ETC1 else [2, 2, 2, 0]
    assert px[0].tolist() == want_px
    err = [255 - v for v in want_px]
    w, h = 4096, 3_200_000
    pixels = w * h
    print(f"ETC quality past one launch: fmt={fmt} {w}x{h}, {pixels // 16} blocks", flush=True)
    src = hip.DeviceBuffer(w * 4 * h, fill=0xFF)
    blocks = hip.DeviceBuffer((w // 4) * (h // 4) * bb, fill=0)
    try:
        q = gpu.quality_device(fmt, blocks, src, w, h, w * 4)
    finally:
        src.free()
        blocks.free()
    assert q["pixels"] == pixels
    assert q["sse_byte"] == [e * e * pixels for e in err]
    y = 21 * err[2] + 72 * err[1] + 7 * err[0]
    assert q["sse_luma"] == pytest.approx(y * y * pixels / 1e4, rel=1e-15)


def test_quality_images_closed_form(gpu):
    from goofy_amd import hip
    w, h = 64, 8
    a = np.tile(np.array([10, 20, 30, 255], np.uint8), w * h)
    b = a.copy()
    b[1::4] += 2  # byte 1 ("g") off by 2 everywhere
    q = gpu.quality_images_device(hip.DeviceBuffer.from_array(a), hip.DeviceBuffer.from_array(b), w, h)
    assert q["sse_byte"] == [0, 4 * w * h, 0, 0]
    assert q["mse_g"] == 4.0 and q["mse_rgb"] == 4.0
    assert q["mse_y"] == pytest.approx((0.72 * 22 - 0.72 * 20) ** 2, rel=1e-12)
    assert q["psnr_rgb"] == pytest.approx(10 * math.log10(768.0 ** 2 / 4.0), abs=1e-12)


def test_quality_calls_from_threads_share_the_scratch_pool(gpu):
    """Quality calls on four threads and streams at once (the library keeps a
    pool of scratch buffers): every result equals the same call made alone."""
    import threading

    from goofy_amd import hip
    rng = np.random.default_rng(3)
    w, h = 256, 64
    srcs, blocks = [], []
    for i in range(4):
        img = rng.integers(0, 256, w * h * 4, dtype=np.uint8)
        srcs.append(hip.DeviceBuffer.from_array(img))
        b = hip.DeviceBuffer((w // 4) * (h // 4) * 8)
        gpu.compress_device(i % 2, b, srcs[-1], w, h, w * 4)
        blocks.append(b)
    hip.synchronize()
    alone = [gpu.quality_device(i % 2, blocks[i], srcs[i], w, h) for i in range(4)]
    got = [[] for _ in range(4)]
    errors = []

    def worker(i):
        try:
            s = hip.Stream()
            for _ in range(20):
                got[i].append(gpu.quality_device(i % 2, blocks[i], srcs[i], w, h, stream=s))
            s.synchronize()
            s.destroy()
        except Exception as e:  # reported below, on the main thread
            errors.append(e)

    print("quality: 4 threads x 20 calls", flush=True)
    threads = [threading.Thread(target=worker, args=(i,)) for i in range(4)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    assert not errors, errors
    for i in range(4):
        assert all(q == alone[i] for q in got[i]), i
