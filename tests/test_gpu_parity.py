"""Parity of the HIP encoder with the reference, on the GPU, through the C-ABI.

Bit-exact on every block: against the golden vectors the reference wrote
(tests/golden/), and against the CPU oracle on the same seeded inputs.
"""
import hashlib
import sys

import numpy as np
import pytest

from goofy_amd import workloads as W

from conftest import REAL_IMAGES

pytestmark = pytest.mark.gpu


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def dev_buffer(nbytes, align=256, fill=None):
    from goofy_amd.hip import DeviceBuffer
    return DeviceBuffer(nbytes, align, fill)


def to_host(buf):
    from goofy_amd import hip
    hip.synchronize()
    return buf.download()


def upload(a, align=256):
    from goofy_amd.hip import DeviceBuffer
    return DeviceBuffer.from_array(np.ascontiguousarray(a).reshape(-1), align)


def encode_dev(g, fmt, src, w, h, pitch, config=None):
    out = dev_buffer(max(8, (w // 4) * (h // 4) * 8), fill=0xCD)
    g.compress_device(fmt, out, src, w, h, pitch, config=config)
    return to_host(out)[: (w // 4) * (h // 4) * 8]


# ---- generator ----------------------------------------------------------------------

@pytest.mark.parametrize("kind,w,h,pitch", [(0, 512, 512, 2048), (1, 300, 132, 1280), (0, 1028, 36, 4112),
                                            (2, 520, 260, 2096)])
def test_synth_device_matches_host(gpu, kind, w, h, pitch):
    say(f"synth kind={kind} {w}x{h} pitch={pitch}")
    host = gpu.synth_host(kind, 99, 4, 17, w, h, pitch, fill=0x5A)
    dev = dev_buffer(pitch * h, fill=0x5A)
    gpu.synth_device(kind, 99, 4, 17, dev, w, h, pitch)
    assert np.array_equal(to_host(dev), host)


# ---- drop-in (host memory) ------------------------------------------------------------

def test_dropin_img512_golden(gpu, golden_img512, digests):
    im = W.IMAGE_512
    img = gpu.synth_host(im.kind, im.seed, im.texture_id, im.noise, im.width, im.height, im.pitch)
    for fmt, key, fn in ((W.DXT1, "dxt1", gpu.compressDXT1), (W.ETC1, "etc1", gpu.compressETC1)):
        say(f"drop-in {key} 512x512")
        out = np.full(im.blocks * 8, 0xCD, np.uint8)
        assert fn(out, img, im.width, im.height, im.pitch) == 0
        assert np.array_equal(out, golden_img512[key])


@pytest.mark.parametrize("w,h,extra", [(16, 4, 0), (32, 8, 16), (48, 100, 4), (528, 36, 200), (1024, 4, 0)])
def test_dropin_strides_vs_oracle(gpu, oracle, w, h, extra):
    rng = np.random.default_rng(w * 7 + h)
    stride = w * 4 + extra
    img = rng.integers(0, 256, stride * h, dtype=np.uint8)  # row padding holds garbage
    for fmt in (W.DXT1, W.ETC1):
        say(f"drop-in fmt={fmt} {w}x{h} stride={stride}")
        rc_o, want = oracle.compress(fmt, img, w, h, stride)
        out = np.full(max(1, (w // 4) * (h // 4) * 8), 0xCD, np.uint8)
        assert gpu.compress_dxt1_fn is not None
        fn = gpu.compressDXT1 if fmt == W.DXT1 else gpu.compressETC1
        assert fn(out, img, w, h, stride) == rc_o == 0
        assert np.array_equal(out[: want.size], want)


def test_dropin_return_codes_on_gpu(gpu):
    out = np.zeros(4096, np.uint8)
    img = np.zeros(64 * 64 * 4, np.uint8)
    assert gpu.compressDXT1(out, img, 60, 64, 240) == -1
    assert gpu.compressETC1(out, img, 64, 62, 256) == -2
    assert gpu.compressETC1(out, img, 0, 8, 0) == 0


# ---- device memory ------------------------------------------------------------------------

def test_device_golden_blocks(gpu, golden_blocks):
    blocks = golden_blocks["blocks"]
    n = blocks.shape[0]
    img = np.ascontiguousarray(blocks.reshape(n, 4, 16).transpose(1, 0, 2)).reshape(-1)  # 4 rows, n blocks wide
    src = upload(img)
    for fmt, key in ((W.DXT1, "dxt1"), (W.ETC1, "etc1")):
        say(f"golden corpus {key}: {n} blocks")
        out = encode_dev(gpu, fmt, src, 4 * n, 4, 16 * n).reshape(n, 8)
        bad = np.nonzero((out != golden_blocks[key]).any(1))[0]
        assert bad.size == 0, f"{bad.size} blocks differ, first {bad[:8]}"


@pytest.mark.parametrize("name", ["waves-8192x8192", "stress-16384x16384-pitch65792"])
def test_device_large_images_golden(gpu, digests, name):
    im = {W.IMAGE_8192.name: W.IMAGE_8192, W.STRESS_16384.name: W.STRESS_16384}[name]
    d = digests["images"][name]
    say(f"device {name}: synth on device")
    src = dev_buffer(im.pitch * im.height, fill=0)
    gpu.synth_device(im.kind, im.seed, im.texture_id, im.noise, src, im.width, im.height, im.pitch)
    assert sha(to_host(src)) == d["input_sha256"]
    for fmt, key in ((W.DXT1, "dxt1"), (W.ETC1, "etc1")):
        say(f"device {name} {key}")
        out = encode_dev(gpu, fmt, src, im.width, im.height, im.pitch)
        assert sha(out) == d[f"{key}_sha256"], key
        # the XCD-contiguous tile order on the same image
        cfg = dict(blocks_per_lane=4, wg_waves=8, waves_per_cu=0, nontemporal=1, tile_order=1)
        say(f"device {name} {key} {cfg}")
        out = encode_dev(gpu, fmt, src, im.width, im.height, im.pitch, cfg)
        assert sha(out) == d[f"{key}_sha256"], (key, cfg)


@pytest.mark.parametrize("seed", range(6))
def test_device_shapes_vs_oracle(gpu, oracle, seed):
    # widths that are multiples of 4 (not 16), padded pitches with poison in
    # the padding, input offsets that are 16- but not 256-byte aligned
    rng = np.random.default_rng(seed)
    w = int(rng.integers(1, 300)) * 4
    h = int(rng.integers(1, 60)) * 4
    pitch = w * 4 + int(rng.integers(0, 5)) * 16
    img = rng.integers(0, 256, pitch * h, dtype=np.uint8)
    if seed % 2:
        img = np.clip(img // 16 + rng.integers(0, 230), 0, 255).astype(np.uint8)
    raw = upload(np.concatenate([np.zeros(48, np.uint8), img]))
    src = raw.view(48, img.nbytes)
    for fmt in (W.DXT1, W.ETC1):
        say(f"shape seed={seed} {w}x{h} pitch={pitch} fmt={fmt}")
        want = oracle.compress_blocks(fmt, img, w, h, pitch)
        assert np.array_equal(encode_dev(gpu, fmt, src, w, h, pitch), want)


@pytest.mark.parametrize("w,h", [(4, 4), (4, 16384), (16384, 4), (65532, 4), (4, 65532), (12, 8)])
def test_extreme_aspect_ratios_vs_oracle(gpu, oracle, w, h):
    # one block column, one block row, a single block, and rows of 16383 blocks:
    # tiles that wrap across many block rows, or never do
    rng = np.random.default_rng(w * 7 + h)
    pitch = w * 4 + 16
    img = rng.integers(0, 256, pitch * h, dtype=np.uint8)
    src = upload(img)
    for fmt in (W.DXT1, W.ETC1):
        say(f"aspect {w}x{h} fmt={fmt}")
        want = oracle.compress_blocks(fmt, img, w, h, pitch)
        assert np.array_equal(encode_dev(gpu, fmt, src, w, h, pitch), want)
        cfg = dict(blocks_per_lane=4, wg_waves=8, waves_per_cu=0, nontemporal=1, tile_order=1)
        assert np.array_equal(encode_dev(gpu, fmt, src, w, h, pitch, cfg), want)


def test_images_beyond_one_launch_are_striped(gpu, oracle):
    """2,097,152 x 16 pixels: 2^19 blocks per row, so one launch covers at most
    3 block rows (block index x blocks per row < 2^40).  Encode, decode and a
    batch entry all run as stripes of 3 + 1 block rows, bit-exact."""
    from goofy_amd import hip
    w, h = 1 << 21, 16
    pitch = w * 4
    rng = np.random.default_rng(21)
    img = rng.integers(0, 256, pitch * h, dtype=np.uint8)
    src = upload(img)
    for fmt in (W.DXT1, W.ETC1):
        say(f"striped {w}x{h} fmt={fmt}")
        want = oracle.compress_blocks(fmt, img, w, h, pitch)
        got_buf = dev_buffer(want.size, fill=0xCD)
        gpu.compress_device(fmt, got_buf, src, w, h, pitch)
        assert np.array_equal(to_host(got_buf), want)
        dec = dev_buffer(pitch * h, fill=0x3C)
        gpu.decode_device(fmt, dec, got_buf, w, h, pitch)
        dec_want = oracle.decode_image(fmt, want, w, h)
        assert np.array_equal(to_host(dec), dec_want)
        # quality sums over the stripes: integer squared errors, and 10^4 x the
        # luma sum as the exact integer sum of (21 dr + 72 dg + 7 db)^2
        d = img.reshape(-1, 4).astype(np.int64) - dec_want.reshape(-1, 4)
        sse = [int(x) for x in (d * d).sum(0)]
        y100 = int(((21 * d[:, 2] + 72 * d[:, 1] + 7 * d[:, 0]) ** 2).sum())
        for q in (gpu.quality_device(fmt, got_buf, src, w, h, pitch), gpu.quality_images_device(src, dec, w, h)):
            assert q["sse_byte"] == sse
            assert q["mse_y"] == pytest.approx(y100 / 1e4 / (w * h), rel=1e-12)
        got_buf.fill(0xCD)
        small = dev_buffer(64 * 64 * 4, fill=7)
        small_out = dev_buffer(16 * 16 * 8, fill=0xCD)
        b = gpu.Batch([dict(src=src, dst=got_buf, width=w, height=h, pitch=pitch, format=fmt),
                       dict(src=small, dst=small_out, width=64, height=64, pitch=256, format=fmt)])
        b.launch()
        hip.synchronize()
        assert np.array_equal(to_host(got_buf), want)
        assert np.array_equal(to_host(small_out),
                              oracle.compress_blocks(fmt, np.full(64 * 64 * 4, 7, np.uint8), 64, 64, 256))
        b.close()
    floor = dev_buffer(pitch * h, fill=0)
    gpu.rgb565_baseline_device(floor, src, w, h)
    assert np.array_equal(to_host(floor), oracle.rgb565_baseline(img, w, h))


def test_non_power_of_two_rows_striped(gpu, oracle):
    """4,194,300 x 16 pixels: 1,048,575 blocks per row (not a power of two, so
    the 40-bit division magic is inexact beyond 2^40 / bw blocks), one block
    row per launch.  Encode, the batch launch, the per-texture launches of a
    batch, decode, the quality sums and the RGB565 floor, bit-exact."""
    from goofy_amd import hip
    w, h = 4 * 1048575, 16
    pitch = w * 4
    rng = np.random.default_rng(22)
    img = rng.integers(0, 256, pitch * h, dtype=np.uint8)
    src = upload(img)
    for fmt in (W.DXT1, W.ETC2):
        bb = 16 if fmt == W.ETC2 else 8
        say(f"non-power-of-two stripes {w}x{h} fmt={fmt}: encode")
        want = oracle.compress_blocks(fmt, img, w, h, pitch)
        got_buf = dev_buffer(want.size, fill=0xCD)
        gpu.compress_device(fmt, got_buf, src, w, h, pitch)
        assert np.array_equal(to_host(got_buf), want)
        say(f"non-power-of-two stripes {w}x{h} fmt={fmt}: batch, per-texture launches")
        for per_texture in (False, True):
            got_buf.fill(0xCD)
            b = gpu.Batch([dict(src=src, dst=got_buf, width=w, height=h, pitch=pitch, format=fmt)])
            b.launch_per_texture() if per_texture else b.launch()
            hip.synchronize()
            assert np.array_equal(to_host(got_buf), want), per_texture
            b.close()
        say(f"non-power-of-two stripes {w}x{h} fmt={fmt}: decode, quality")
        dec = dev_buffer(pitch * h, fill=0x3C)
        gpu.decode_device(fmt, dec, got_buf, w, h, pitch)
        dec_want = oracle.decode_image(fmt, want, w, h)
        assert np.array_equal(to_host(dec), dec_want)
        d = img.reshape(-1, 4).astype(np.int64) - dec_want.reshape(-1, 4)
        assert gpu.quality_device(fmt, got_buf, src, w, h, pitch)["sse_byte"] == [int(x) for x in (d * d).sum(0)]
        assert want.size == (w // 4) * (h // 4) * bb
        del dec
    say(f"non-power-of-two stripes {w}x{h}: rgb565 floor")
    floor = dev_buffer(pitch * h, fill=0)
    gpu.rgb565_baseline_device(floor, src, w, h)
    assert np.array_equal(to_host(floor), oracle.rgb565_baseline(img, w, h))


def test_65536_wide_image_striped_vs_reference(gpu, reference):
    """65536 x 16384 pixels (4 GiB, generated on the device): blocks x blocks per
    row is just above 2^40, so the encode runs as 2 stripes.  Both formats,
    all 67M blocks against the reference encoder."""
    from goofy_amd import hip
    from oracle.oracle import aligned_empty
    w, h = 65536, 16384
    pitch = w * 4
    nb = (w // 4) * (h // 4)
    say(f"wide striped {w}x{h}: synth on device")
    src = dev_buffer(pitch * h)
    gpu.synth_device(0, 0x601F000A, 0, 12, src, w, h, pitch)
    img = aligned_empty(pitch * h)
    hip.check(hip._hip.hipMemcpy(img.ctypes.data, src.data_ptr(), pitch * h, hip.D2H), "download")
    out = dev_buffer(nb * 8, fill=0xCD)
    for fmt in (W.DXT1, W.ETC1):
        say(f"wide striped {w}x{h} fmt={fmt}")
        gpu.compress_device(fmt, out, src, w, h, pitch)
        want = reference.compress_mt(fmt, img, w, h, pitch, 16)
        assert sha(to_host(out)) == sha(want), fmt
    del img


# ---- both formats in one pass ----------------------------------------------------------------

@pytest.mark.parametrize("name", ["waves-8192x8192", "stress-16384x16384-pitch65792"])
def test_dual_golden_digests(gpu, digests, name):
    im = {W.IMAGE_8192.name: W.IMAGE_8192, W.STRESS_16384.name: W.STRESS_16384}[name]
    d = digests["images"][name]
    src = dev_buffer(im.pitch * im.height, fill=0)
    gpu.synth_device(im.kind, im.seed, im.texture_id, im.noise, src, im.width, im.height, im.pitch)
    a = dev_buffer(im.blocks * 8, fill=0xCD)
    b = dev_buffer(im.blocks * 8, fill=0xCD)
    say(f"dual {name}")
    gpu.compress_device_dual(a, b, src, im.width, im.height, im.pitch)
    assert sha(to_host(a)) == d["dxt1_sha256"]
    assert sha(to_host(b)) == d["etc1_sha256"]


@pytest.mark.parametrize("seed", range(4))
def test_dual_shapes_vs_oracle(gpu, oracle, seed):
    rng = np.random.default_rng(100 + seed)
    w = int(rng.integers(1, 400)) * 4
    h = int(rng.integers(1, 80)) * 4
    pitch = w * 4 + int(rng.integers(0, 4)) * 16
    img = rng.integers(0, 256, pitch * h, dtype=np.uint8)
    if seed % 2:
        img = np.clip(img // 8 + rng.integers(0, 220), 0, 255).astype(np.uint8)
    src = upload(img)
    n = (w // 4) * (h // 4) * 8
    a, b = dev_buffer(n, fill=0xCD), dev_buffer(n, fill=0xCD)
    say(f"dual shape seed={seed} {w}x{h} pitch={pitch}")
    gpu.compress_device_dual(a, b, src, w, h, pitch)
    assert np.array_equal(to_host(a), oracle.compress_blocks(W.DXT1, img, w, h, pitch))
    assert np.array_equal(to_host(b), oracle.compress_blocks(W.ETC1, img, w, h, pitch))


def test_dual_striped_and_random_blocks(gpu, reference):
    """The dual kernel on 2^19-block rows (stripes of 3 + 1 block rows) of
    threshold-kind random blocks, against the reference encoder."""
    from blockgen import block_image
    bw, bh = 1 << 19, 4
    w, h = 4 * bw, 4 * bh
    img = block_image("threshold", 77, bw, bh)
    src = upload(img)
    n = bw * bh * 8
    a, b = dev_buffer(n, fill=0xCD), dev_buffer(n, fill=0xCD)
    say(f"dual striped {w}x{h}")
    gpu.compress_device_dual(a, b, src, w, h, w * 4)
    assert np.array_equal(to_host(a), reference.compress_mt(W.DXT1, img, w, h, w * 4, 16))
    assert np.array_equal(to_host(b), reference.compress_mt(W.ETC1, img, w, h, w * 4, 16))


@pytest.mark.parametrize("bpl", [1, 2, 4])
@pytest.mark.parametrize("wg", [1, 2, 4, 8])
@pytest.mark.parametrize("nt", [0, 1])
@pytest.mark.parametrize("order", [0, 1])
def test_launch_configs_vs_oracle(gpu, oracle, bpl, wg, nt, order):
    # 257 x 65 blocks: the tile count is not a multiple of any workgroup size
    # or of the 8 XCDs, so every remapped launch has a partial last share
    w, h, pitch = 1028, 260, 1028 * 4 + 48
    img = gpu.synth_host(0, 1234, 0, 25, w, h, pitch, fill=0x77)
    src = upload(img)
    cfg = dict(blocks_per_lane=bpl, wg_waves=wg, waves_per_cu=0, nontemporal=nt, tile_order=order)
    for fmt in (W.DXT1, W.ETC1):
        say(f"config {cfg} fmt={fmt}")
        assert np.array_equal(encode_dev(gpu, fmt, src, w, h, pitch, cfg), oracle.compress_blocks(fmt, img, w, h, pitch))


def test_tune_device_picks_a_valid_shape(gpu, oracle):
    w, h, pitch = 2048, 1024, 2048 * 4 + 256
    img = gpu.synth_host(0, 99, 0, 12, w, h, pitch, fill=0)
    src = upload(img)
    out = dev_buffer((w // 4) * (h // 4) * 8, fill=0xCD)
    cfg, ms = gpu.tune_device(W.ETC1, out, src, w, h, pitch, reps=2)
    say(f"tuned {cfg} {ms:.4f} ms")
    want = oracle.compress_blocks(W.ETC1, img, w, h, pitch)
    assert ms > 0 and cfg["blocks_per_lane"] in (1, 2, 4) and cfg["tile_order"] in (0, 1)
    assert np.array_equal(to_host(out), want)  # the tuner leaves a valid encoding behind
    out.fill(0xCD)
    assert np.array_equal(encode_dev(gpu, W.ETC1, src, w, h, pitch, cfg), want)


# ---- batched --------------------------------------------------------------------------------

def make_batch_inputs(g, textures):
    items = []
    for t in textures:
        src = dev_buffer(t["width"] * 4 * t["height"])
        g.synth_device(t["kind"], t["seed"], t["texture_id"], t["noise"], src, t["width"], t["height"])
        dst = dev_buffer((t["width"] // 4) * (t["height"] // 4) * 8, fill=0xCD)
        items.append(dict(src=src, dst=dst, width=t["width"], height=t["height"], pitch=t["width"] * 4,
                          format=t["format"], meta=t))
    return items


def test_batch_golden16(gpu, digests):
    ts = digests["batch"]
    items = make_batch_inputs(gpu, ts)
    say(f"batch golden: {len(items)} textures")
    b = gpu.Batch(items)
    b.launch()
    for it, t in zip(items, ts):
        assert sha(to_host(it["src"])) == t["input_sha256"]
        assert sha(to_host(it["dst"])) == t["output_sha256"], t


def test_batch_vs_oracle_and_configs(gpu, oracle):
    from goofy_amd import hip
    ts = W.batch_textures(n=40, lo=4, hi=700, seed=77)
    items = make_batch_inputs(gpu, ts)
    want = []
    for it in items:
        img = to_host(it["src"])
        want.append(oracle.compress_blocks(it["format"], img, it["width"], it["height"], it["pitch"]))
    b = gpu.Batch(items)
    say(f"batch 40 textures, info={b.info()}, config={b.config}")

    def check(tag):
        for it, wv in zip(items, want):
            assert np.array_equal(to_host(it["dst"]), wv), (tag, it["meta"])

    def clear():
        for it in items:
            it["dst"].fill(0xCD)

    b.launch()
    check("graph")
    # replays of the captured graph, with eager device work in between
    scratch = dev_buffer(1 << 20)
    for r in range(3):
        clear()
        hip.synchronize()
        # eager device work between replays
        gpu.synth_device(0, r, 0, 5, scratch, 256, 1024)
        b.launch()
        hip.synchronize()
        assert scratch.download(16).any()
        check(f"replay {r}")
    clear()
    b.launch_per_texture()
    check("per-texture")
    for cfg in (dict(blocks_per_lane=1, wg_waves=2, waves_per_cu=0, nontemporal=0),
                dict(blocks_per_lane=4, wg_waves=8, waves_per_cu=1, nontemporal=1),
                dict(blocks_per_lane=2, wg_waves=1, waves_per_cu=3, nontemporal=0),
                dict(blocks_per_lane=1, wg_waves=2, waves_per_cu=0, nontemporal=1, tile_order=1),
                dict(blocks_per_lane=2, wg_waves=4, waves_per_cu=5, nontemporal=1, tile_order=1)):
        clear()
        b.config = cfg
        say(f"batch config {cfg}")
        b.launch()
        check(str(cfg))
    clear()
    chosen, ms = b.autotune(reps=1)
    say(f"autotune chose {chosen} ({ms:.3f} ms)")
    clear()
    b.launch()
    check("autotuned")
    b.close()


def test_batch_full_workload(gpu, oracle):
    """The 4096-texture workload: batch output equals per-texture launches on
    every texture, and the oracle on a sample of textures."""
    from goofy_amd import hip
    ts = W.batch_textures()
    say("batch 4096: generating on device")
    items = make_batch_inputs(gpu, ts)
    b = gpu.Batch(items)
    say(f"batch 4096 info={b.info()}")
    b.launch()
    hip.synchronize()
    first = [it["dst"].download() for it in items]
    for it in items:
        it["dst"].fill(0xCD)
    b.launch_per_texture()
    hip.synchronize()
    for f, it in zip(first, items):
        assert np.array_equal(f, it["dst"].download()), it["meta"]
    rng = np.random.default_rng(0)
    for i in rng.choice(len(items), 48, replace=False):
        it = items[int(i)]
        img = to_host(it["src"])
        want = oracle.compress_blocks(it["format"], img, it["width"], it["height"], it["pitch"])
        assert np.array_equal(first[int(i)], want), it["meta"]
    b.close()


# ---- real images (the reference's own test data) -------------------------------------------

def _real():
    import json
    from conftest import GOLDEN
    return GOLDEN, json.loads((GOLDEN / "real_images.json").read_text())


@pytest.mark.parametrize("name", REAL_IMAGES)
def test_real_images_gpu(gpu, name):
    from goofy_amd import io as gio
    golden, real = _real()
    g = real[name]
    img, w, h = gio.load_png_rgba8(golden / "images" / name)
    src = upload(img)
    for fmt, key, fn in ((W.DXT1, "dxt1", gpu.compressDXT1), (W.ETC1, "etc1", gpu.compressETC1)):
        say(f"real image {name} {key}")
        assert sha(encode_dev(gpu, fmt, src, w, h, w * 4)) == g[f"{key}_sha256"]
        out = np.zeros((w // 4) * (h // 4) * 8, np.uint8)
        assert fn(out, img, w, h, w * 4) == 0 and sha(out) == g[f"{key}_sha256"]
        blocks = upload(out)
        q = gpu.quality_device(fmt, blocks, src, w, h, w * 4)
        assert q["sse_byte"] == g[f"{key}_psnr"]["sse_byte"]
        assert q["psnr_rgb"] == pytest.approx(g[f"{key}_psnr"]["psnr_rgb"], abs=1e-9)
        assert q["psnr_y"] == pytest.approx(g[f"{key}_psnr"]["psnr_y"], abs=1e-7)


@pytest.mark.parametrize("name", ["parrot_red.png", "roblox01.png", "patterns.png"])
def test_containers_of_gpu_blocks_match_harness(gpu, reference, oracle, harness, tmp_path, name):
    """DDS and KTX1 files written through goofy_container_write around blocks
    the GPU encoded from a real image, byte for byte against the files the
    harness's saveDds / saveKtx (Src/main.cpp:154-220) write around the
    reference's blocks (DXT1, ETC1S) or the oracle's (DXT5, ETC2 RGBA8: the
    reference has no alpha encoder)."""
    from goofy_amd import io as gio
    golden, real = _real()
    img, w, h = gio.load_png_rgba8(golden / "images" / name)
    src = upload(img)
    for fmt in (W.DXT1, W.ETC1, W.DXT5, W.ETC2):
        bb = gpu.block_bytes(fmt)
        out = dev_buffer((w // 4) * (h // 4) * bb, fill=0xCD)
        gpu.compress_device(fmt, out, src, w, h, w * 4)
        got = to_host(out)
        if fmt in (W.DXT1, W.ETC1):
            rc, want = reference.compress(fmt, img, w, h, w * 4)
            assert rc == 0
        else:
            want = oracle.compress_blocks(fmt, img, w, h, w * 4)
        for kind in (gio.DDS, gio.KTX):
            if kind == gio.DDS and fmt in (W.ETC1, W.ETC2):
                continue  # the harness writes ETC formats as KTX only
            say(f"container {name} fmt={fmt} kind={kind}")
            path = tmp_path / f"{kind}_{fmt}"
            harness.save(kind, fmt, path, want, w, h)
            assert gio.container_bytes(kind, fmt, w, h, got) == path.read_bytes(), (name, fmt, kind)


@pytest.mark.parametrize("name", ["parrot_red.png", "kodim23.png", "roblox01.png"])
def test_harness_plugin_loop_runs_hip_entry_points(gpu, harness, tmp_path, monkeypatch, name):
    """The reference harness's own runTestDXT1 / runTestETC1
    (Src/main.cpp:646-743) with goofy_hip_compress_dxt1 / _etc1 in its
    CompressFunc_t slot (:644), beside the harness's compiled
    goofy::compressDXT1 / compressETC1 in the same loop: identical blocks,
    getMsePsnr figures and .dds / .ktx / .tga files; the best-of-128 times of
    both are printed."""
    import ctypes

    from conftest import GOLDEN
    monkeypatch.chdir(tmp_path)
    (tmp_path / "test-results").mkdir()
    img, w, h = harness.load_png(GOLDEN / "images" / name)  # the harness's own loader
    stem = name[:-4]
    for fmt, key, ext, fn in ((W.DXT1, "dxt1", ".dds", gpu._lib.goofy_hip_compress_dxt1),
                              (W.ETC1, "etc1", ".ktx", gpu._lib.goofy_hip_compress_etc1)):
        say(f"harness loop {name} {key}")
        addr = ctypes.cast(fn, ctypes.c_void_p).value
        ref_blocks, ref_us, ref_q = harness.run_test(fmt, "simd_goofy", stem, harness.goofy_encoder(fmt), 128, img,
                                                     w, h, w * 4)
        blocks, us, q = harness.run_test(fmt, "hip_goofy", stem, addr, 128, img, w, h, w * 4)
        assert np.array_equal(blocks, ref_blocks)
        assert q == ref_q
        for suffix in (f"_{key}{ext}", f"_{key}_decompressed.tga"):
            a = (tmp_path / "test-results" / f"{stem}_hip_goofy{suffix}").read_bytes()
            assert a == (tmp_path / "test-results" / f"{stem}_simd_goofy{suffix}").read_bytes(), suffix
        say(f"harness loop {name} {key}: psnrRGB {q['psnr_rgb']:.3f} dB; best of 128: reference {ref_us:.0f} us "
            f"(1 thread), HIP drop-in {us:.0f} us")


def test_device_image_beyond_4gib(gpu, oracle):
    """32768x32768 RGBA8 (4 GiB of pixels, 67M blocks): byte offsets past 2^32.
    Stripes of block rows at the start, the middle and the very end are compared
    with the oracle (blocks are independent, so a stripe's blocks are exactly
    the encoding of that stripe as an image)."""
    from goofy_amd import hip
    w = h = 32768
    pitch = w * 4
    say("device 32768x32768 stress: synth on device (4 GiB)")
    src = dev_buffer(pitch * h)
    gpu.synth_device(1, 0x601F0099, 0, 0, src, w, h, pitch)
    out = dev_buffer((w // 4) * (h // 4) * 8)
    for fmt in (W.DXT1, W.ETC1):
        say(f"device 32768x32768 fmt={fmt}")
        gpu.compress_device(fmt, out, src, w, h, pitch)
        hip.synchronize()
        row_blocks = (w // 4) * 8
        for by in (0, 4097, 6000, h // 4 - 2):
            rows = 2
            img = src.download(pitch * 4 * rows, offset=pitch * 4 * by)
            want = oracle.compress_blocks(fmt, img, w, 4 * rows, pitch)
            got = out.download(row_blocks * rows, offset=row_blocks * by)
            assert np.array_equal(got, want), (fmt, by)
    src.free()
    out.free()


# ---- random blocks against the reference encoder itself ---------------------------------------

@pytest.mark.parametrize("kind", ["uniform", "narrow", "ramps", "threshold"])
def test_random_blocks_vs_reference(gpu, reference, kind):
    """2,621,440 random blocks per kind (10.5M over the four kinds, see
    tests/blockgen.py), both formats, bit-exact against the reference encoder
    compiled from its own source (oracle/_ref)."""
    from blockgen import block_image
    bw, bh = 4096, 640
    w, h = 4 * bw, 4 * bh
    img = block_image(kind, {"uniform": 1, "narrow": 2, "ramps": 3, "threshold": 4}[kind], bw, bh)
    src = upload(img)
    for fmt in (W.DXT1, W.ETC1):
        say(f"random blocks {kind} fmt={fmt}: {bw * bh} blocks")
        got = encode_dev(gpu, fmt, src, w, h, w * 4)
        want = reference.compress_mt(fmt, img, w, h, w * 4, 16)
        bad = np.nonzero((got.reshape(-1, 8) != want.reshape(-1, 8)).any(1))[0]
        assert bad.size == 0, f"{bad.size} blocks differ, first {bad[:8]}"


@pytest.mark.parametrize("w,h,extra", [(16, 4, 0), (528, 36, 200), (1024, 4100, 12), (4096, 1028, 0)])
def test_dropin_dual_matches_single_format(gpu, oracle, w, h, extra):
    """goofy_hip_compress_dual against the two single-format drop-ins; the
    last two shapes span several ~8 MiB staging chunks, the last chunk short."""
    rng = np.random.default_rng(w + 3 * h)
    stride = w * 4 + extra
    img = rng.integers(0, 256, stride * h, dtype=np.uint8)
    img[::7] = img[::7] // 8  # some narrow-range blocks too
    say(f"drop-in dual {w}x{h} stride={stride}")
    a, b = gpu.compress_dual(img, w, h, stride)
    n = (w // 4) * (h // 4) * 8
    for fmt, got, fn in ((W.DXT1, a, gpu.compressDXT1), (W.ETC1, b, gpu.compressETC1)):
        one = np.full(max(1, n), 0xCD, np.uint8)
        assert fn(one, img, w, h, stride) == 0
        assert np.array_equal(got, one[:n])
        if n <= 1 << 16:
            rc, want = oracle.compress(fmt, img, w, h, stride)
            assert rc == 0 and np.array_equal(got, want)


def test_dropin_dual_return_codes(gpu):
    img = np.zeros(64 * 64 * 4, np.uint8)
    out_a, out_b = np.zeros(4096, np.uint8), np.zeros(4096, np.uint8)
    f = gpu._lib.goofy_hip_compress_dual
    assert f(gpu._u8(out_a), gpu._u8(out_b), gpu._u8(img), 60, 64, 240) == -1
    assert f(gpu._u8(out_a), gpu._u8(out_b), gpu._u8(img), 64, 62, 256) == -2
    assert f(gpu._u8(out_a), gpu._u8(out_b), gpu._u8(img), 0, 8, 0) == 0
    assert f(gpu._u8(out_a), gpu._u8(None), gpu._u8(img), 64, 64, 256) == gpu.ERR_INVALID


@pytest.mark.parametrize("w,h,extra", [(2048, 8192, 0), (4096, 4100, 48)])
def test_dropin_many_chunks_matches_device(gpu, w, h, extra):
    """Host images of 8 and 16 staging chunks (every slot reused several times,
    the last chunk short): the drop-in against the device entry point."""
    stride = w * 4 + extra
    img = gpu.synth_host(0, 4242, 1, 20, w, h, stride, fill=0x11)
    src = upload(img)
    n = (w // 4) * (h // 4) * 8
    for fmt, fn in ((W.DXT1, gpu.compressDXT1), (W.ETC1, gpu.compressETC1)):
        say(f"drop-in many chunks fmt={fmt} {w}x{h} stride={stride}")
        want = encode_dev(gpu, fmt, src, w, h, stride)
        out = np.full(n, 0xCD, np.uint8)
        assert fn(out, img, w, h, stride) == 0
        assert np.array_equal(out, want)
    a, b = gpu.compress_dual(img, w, h, stride)
    assert np.array_equal(a, encode_dev(gpu, W.DXT1, src, w, h, stride))
    assert np.array_equal(b, encode_dev(gpu, W.ETC1, src, w, h, stride))


@pytest.mark.parametrize("w,h,extra", [(528, 1000, 16), (1040, 1004, 48), (2032, 1028, 0)])
def test_dropin_zero_copy_parts_match_device(gpu, w, h, extra):
    """Host images of one staging chunk (at most about 8 MiB) run zero-copy,
    staged and launched in up to 8 parts of at least 512 KiB.  These shapes
    split into 4, 7 and 8 parts whose block rows do not divide evenly (the last
    part short), one with padded rows: every format, and both pairs from one
    call, against the device entry point."""
    stride = w * 4 + extra
    img = gpu.synth_host(2, 777, 3, 16, w, h, stride, fill=0x5A)  # waves with alpha
    src = upload(img)
    nb = (w // 4) * (h // 4)
    for fmt, fn in ((W.DXT1, gpu.compressDXT1), (W.ETC1, gpu.compressETC1), (W.DXT5, gpu.compressDXT5),
                    (W.ETC2, gpu.compressETC2)):
        say(f"drop-in zero-copy parts fmt={fmt} {w}x{h} stride={stride}")
        bb = gpu.block_bytes(fmt)
        dev = dev_buffer(nb * bb, fill=0xCD)
        gpu.compress_device(fmt, dev, src, w, h, stride)
        want = to_host(dev)[: nb * bb]
        out = np.full(nb * bb, 0xCD, np.uint8)
        assert fn(out, img, w, h, stride) == 0
        assert np.array_equal(out, want)
    a, b = gpu.compress_dual(img, w, h, stride)
    assert np.array_equal(a, encode_dev(gpu, W.DXT1, src, w, h, stride))
    assert np.array_equal(b, encode_dev(gpu, W.ETC1, src, w, h, stride))
    c, d = gpu.compress_dual_alpha(img, w, h, stride)
    for fmt, got in ((W.DXT5, c), (W.ETC2, d)):
        dev = dev_buffer(nb * 16, fill=0xCD)
        gpu.compress_device(fmt, dev, src, w, h, stride)
        assert np.array_equal(got, to_host(dev)[: nb * 16])


def test_dropin_from_threads(gpu):
    """The drop-in on four host threads at once, each with its own images of
    one to several staging chunks: every call bit-identical to the same call
    made alone."""
    import threading
    shapes = [(512, 512), (1024, 2052), (2048, 3072), (4096, 1028)]
    imgs = [gpu.synth_host(0, 77 + i, i, 15, w, h) for i, (w, h) in enumerate(shapes)]
    fmts = [W.DXT1, W.ETC1, W.DXT5, W.ETC2]
    alone = {}
    for i, (w, h) in enumerate(shapes):
        for f in fmts:
            alone[i, f] = gpu.compress(f, imgs[i], w, h)
    errors = []

    def worker(i):
        w, h = shapes[i]
        try:
            for rep in range(6):
                for f in fmts:
                    if not np.array_equal(gpu.compress(f, imgs[i], w, h), alone[i, f]):
                        errors.append((i, f, rep))
        except Exception as e:  # reported below, on the main thread
            errors.append(repr(e))

    say("drop-in: 4 threads x 24 calls")
    threads = [threading.Thread(target=worker, args=(i,)) for i in range(4)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    assert not errors, errors[:5]
