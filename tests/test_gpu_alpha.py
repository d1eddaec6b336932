"""DXT5 (BC3) and ETC2 RGBA8 (EAC alpha) on the GPU, through the C-ABI.

Bit-exact against the CPU oracle, whose alpha rule the CPU tests pin through
the reference's own decoders (tests/test_alpha.py); the colour half of every
16-byte block is also checked against the GPU's own DXT1 / ETC1S output.
"""
import sys

import numpy as np
import pytest

from goofy_amd import workloads as W

pytestmark = pytest.mark.gpu

ALPHA = [(W.DXT5, W.DXT1), (W.ETC2, W.ETC1)]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def dev_buffer(nbytes, fill=None):
    from goofy_amd.hip import DeviceBuffer
    return DeviceBuffer(nbytes, 256, fill)


def to_host(buf):
    from goofy_amd import hip
    hip.synchronize()
    return buf.download()


def upload(a):
    from goofy_amd.hip import DeviceBuffer
    return DeviceBuffer.from_array(np.ascontiguousarray(a).reshape(-1), 256)


def encode_dev(g, fmt, src, w, h, pitch, config=None):
    n = (w // 4) * (h // 4) * g.block_bytes(fmt)
    out = dev_buffer(max(16, n), fill=0xCD)
    g.compress_device(fmt, out, src, w, h, pitch, config=config)
    return to_host(out)[:n]


@pytest.mark.parametrize("fmt,cfmt", ALPHA)
@pytest.mark.parametrize("name", ["waves-alpha-512x512", "waves-alpha-8192x8192"])
def test_alpha_images_vs_oracle(gpu, oracle, name, fmt, cfmt):
    im = {W.ALPHA_512.name: W.ALPHA_512, W.ALPHA_8192.name: W.ALPHA_8192}[name]
    img = gpu.synth_host(im.kind, im.seed, im.texture_id, im.noise, im.width, im.height, im.pitch)
    a = img.reshape(-1, 4)[:, 3]
    assert (a == 255).any() and (a == 0).any() and ((a > 0) & (a < 255)).mean() > 0.2  # all three region kinds
    src = dev_buffer(im.pitch * im.height, fill=0)
    gpu.synth_device(im.kind, im.seed, im.texture_id, im.noise, src, im.width, im.height, im.pitch)
    say(f"alpha {name} fmt={fmt}")
    got = encode_dev(gpu, fmt, src, im.width, im.height, im.pitch)
    want = oracle.compress_blocks(fmt, img, im.width, im.height, im.pitch)
    assert np.array_equal(got, want)
    colour = encode_dev(gpu, cfmt, src, im.width, im.height, im.pitch)
    assert np.array_equal(got.reshape(-1, 16)[:, 8:].reshape(-1), colour)


@pytest.mark.parametrize("fmt", [W.DXT5, W.ETC2])
@pytest.mark.parametrize("kind", ["alpha", "uniform", "threshold"])
def test_alpha_random_blocks_vs_oracle(gpu, oracle, fmt, kind):
    from blockgen import block_image
    bw, bh = 1024, 256  # 262,144 blocks
    img = block_image(kind, 31 + fmt, bw, bh)
    src = upload(img)
    say(f"alpha random blocks {kind} fmt={fmt}")
    got = encode_dev(gpu, fmt, src, 4 * bw, 4 * bh, 16 * bw)
    assert np.array_equal(got, oracle.compress_blocks(fmt, img, 4 * bw, 4 * bh, 16 * bw))


@pytest.mark.parametrize("bpl,wg,nt,order", [(1, 8, 1, 0), (2, 4, 0, 1), (4, 1, 1, 1), (1, 2, 0, 0)])
def test_alpha_launch_configs_vs_oracle(gpu, oracle, bpl, wg, nt, order):
    w, h, pitch = 1028, 260, 1028 * 4 + 48
    img = gpu.synth_host(W.WAVES_ALPHA, 4321, 0, 25, w, h, pitch, fill=0x77)
    src = upload(img)
    cfg = dict(blocks_per_lane=bpl, wg_waves=wg, waves_per_cu=0, nontemporal=nt, tile_order=order)
    for fmt in (W.DXT5, W.ETC2):
        say(f"alpha config {cfg} fmt={fmt}")
        assert np.array_equal(encode_dev(gpu, fmt, src, w, h, pitch, cfg),
                              oracle.compress_blocks(fmt, img, w, h, pitch))


@pytest.mark.parametrize("bpl", [1, 2, 4])
@pytest.mark.parametrize("w", [400, 200])
def test_narrow_textures_all_formats_vs_oracle(gpu, oracle, bpl, w):
    # 100 and 50 blocks a row: below 64 * blocks-per-lane a tile spans more
    # than two block rows and every lane forms its own 64-bit address; at 100
    # blocks and one block a lane, tiles straddle two rows on the wave-uniform
    # base path
    h, pitch = 84, w * 4 + 32
    img = gpu.synth_host(W.WAVES_ALPHA, 77 + w, 0, 25, w, h, pitch, fill=0x3C)
    src = upload(img)
    cfg = dict(blocks_per_lane=bpl, wg_waves=4, waves_per_cu=0, nontemporal=1, tile_order=1)
    for fmt in (W.DXT1, W.ETC1, W.DXT5, W.ETC2):
        say(f"narrow {w}x{h} config {cfg} fmt={fmt}")
        assert np.array_equal(encode_dev(gpu, fmt, src, w, h, pitch, cfg),
                              oracle.compress_blocks(fmt, img, w, h, pitch))


@pytest.mark.parametrize("pitch", [(1 << 29) - 16, 1 << 29])
def test_pitch_either_side_of_the_32_bit_offset_limit(gpu, oracle, pitch):
    # A pixel's byte offset from its tile's first block row reaches
    # 7 * pitch + 16 * bw; below pitch = 2^29 the kernels address a tile from
    # one wave-uniform base with 32-bit offsets (here up to 3.7 GB), from there
    # on with per-lane 64-bit addresses.  68 blocks a row: tiles straddle the
    # two block rows.  About 3.7 GB of device memory.
    w, h, tight = 272, 8, 272 * 4
    img = gpu.synth_host(W.WAVES_ALPHA, 31, 0, 25, w, h, tight, fill=0)
    big = dev_buffer(pitch * (h - 1) + tight)
    gpu.synth_device(W.WAVES_ALPHA, 31, 0, 25, big, w, h, pitch)
    try:
        for bpl in (1, 2):
            cfg = dict(blocks_per_lane=bpl, wg_waves=2, waves_per_cu=0, nontemporal=1, tile_order=0)
            for fmt in (W.DXT1, W.ETC1, W.DXT5, W.ETC2):
                say(f"pitch {pitch:#x} bpl={bpl} fmt={fmt}")
                assert np.array_equal(encode_dev(gpu, fmt, big, w, h, pitch, cfg),
                                      oracle.compress_blocks(fmt, img, w, h, tight))
        for fn, fa, fb in ((gpu.compress_device_dual, W.DXT1, W.ETC1),
                           (gpu.compress_device_dual_alpha, W.DXT5, W.ETC2)):
            na, nb = (w // 4) * (h // 4) * gpu.block_bytes(fa), (w // 4) * (h // 4) * gpu.block_bytes(fb)
            a, b = dev_buffer(na, fill=0xCD), dev_buffer(nb, fill=0xCD)
            say(f"pitch {pitch:#x} {fn.__name__}")
            fn(a, b, big, w, h, pitch)
            assert np.array_equal(to_host(a)[:na], oracle.compress_blocks(fa, img, w, h, tight))
            assert np.array_equal(to_host(b)[:nb], oracle.compress_blocks(fb, img, w, h, tight))
    finally:
        big.free()


def test_alpha_striped(gpu, oracle):
    """2^19-block rows: the image runs as stripes of 3 + 1 block rows."""
    from blockgen import block_image
    bw, bh = 1 << 19, 4
    img = block_image("alpha", 78, bw, bh)
    src = upload(img)
    for fmt in (W.DXT5, W.ETC2):
        say(f"alpha striped {4 * bw}x{4 * bh} fmt={fmt}")
        got = encode_dev(gpu, fmt, src, 4 * bw, 4 * bh, 16 * bw)
        assert np.array_equal(got, oracle.compress_blocks(fmt, img, 4 * bw, 4 * bh, 16 * bw))


@pytest.mark.parametrize("w,h,extra", [(16, 4, 0), (48, 100, 4), (528, 36, 200), (2048, 1028, 16)])
def test_alpha_dropin_vs_oracle(gpu, oracle, w, h, extra):
    rng = np.random.default_rng(w + h)
    stride = w * 4 + extra
    img = rng.integers(0, 256, stride * h, dtype=np.uint8)
    img[3::8] = 255  # half the pixels opaque
    n = (w // 4) * (h // 4) * 16
    for fmt, fn, cfn in ((W.DXT5, gpu.compressDXT5, gpu.compress_dxt5_fn),
                         (W.ETC2, gpu.compressETC2, gpu.compress_etc2_fn)):
        say(f"alpha drop-in fmt={fmt} {w}x{h} stride={stride}")
        out = np.full(n, 0xCD, np.uint8)
        assert fn(out, img, w, h, stride) == 0
        rc, want = oracle.compress(fmt, img, w, h, stride)
        assert rc == 0 and np.array_equal(out, want)
        assert np.array_equal(gpu.compress(fmt, img, w, h, stride), want)
        assert fn(out, img, w - 4, h, stride) == -1 and fn(out, img, w, h - 2, stride) == -2


def mixed_textures(n, seed, lo, hi):
    """Textures of all four formats (alpha content), sizes log-uniform multiples of 4."""
    rng = np.random.default_rng(seed)
    ts = W.batch_textures(n=n, lo=lo, hi=hi, seed=seed)
    fmts = rng.integers(0, 4, n)
    return [dict(t, kind=W.WAVES_ALPHA, format=int(f)) for t, f in zip(ts, fmts)]


def test_alpha_batch_all_formats(gpu, oracle):
    from goofy_amd import hip
    ts = mixed_textures(40, 78, 4, 600)
    items, want = [], []
    for t in ts:
        src = dev_buffer(t["width"] * 4 * t["height"])
        gpu.synth_device(t["kind"], t["seed"], t["texture_id"], t["noise"], src, t["width"], t["height"])
        n = (t["width"] // 4) * (t["height"] // 4) * gpu.block_bytes(t["format"])
        items.append(dict(src=src, dst=dev_buffer(n, fill=0xCD), width=t["width"], height=t["height"],
                          pitch=t["width"] * 4, format=t["format"]))
        img = gpu.synth_host(t["kind"], t["seed"], t["texture_id"], t["noise"], t["width"], t["height"])
        want.append(oracle.compress_blocks(t["format"], img, t["width"], t["height"], t["width"] * 4))
    assert {t["format"] for t in ts} == {0, 1, 2, 3}
    b = gpu.Batch(items)

    def check(tag):
        for it, wv in zip(items, want):
            assert np.array_equal(to_host(it["dst"]), wv), (tag, it["format"], it["width"], it["height"])

    def clear():
        for it in items:
            it["dst"].fill(0xCD)

    say(f"alpha batch 40 textures, all formats, config={b.config}")
    b.launch()
    check("graph")
    for r in range(2):
        clear()
        hip.synchronize()
        b.launch()
        check(f"replay {r}")
    clear()
    b.launch_per_texture()
    check("per-texture")
    for cfg in (dict(blocks_per_lane=2, wg_waves=4, waves_per_cu=0, nontemporal=1, tile_order=1),
                dict(blocks_per_lane=4, wg_waves=8, waves_per_cu=2, nontemporal=0, tile_order=0)):
        clear()
        b.config = cfg
        say(f"alpha batch config {cfg}")
        b.launch()
        check(str(cfg))
    b.close()


def test_alpha_serving_all_formats(gpu):
    from goofy_amd import hip, serving
    ts = mixed_textures(32, 79, 4, 640)
    pool, expected, keep = serving.make_pool(gpu, hip, ts)
    say("alpha serving: Poisson 20k req/s, 0.3 s, every result verified")
    res = serving.simulate(pool, 20000.0, 0.3, seed=4, expected=expected, verify_every=1)
    say(res)
    keep.free()
    assert res["completed"] == res["submitted"] > 2000
    assert res["verified"] == res["completed"] and res["mismatches"] == 0


def test_alpha_large_fuzz_vs_oracle(gpu, oracle):
    """4,194,304 random blocks (alpha kind: six alpha distributions per block,
    random colour) in one 8192 x 2048 image, both 16-byte formats against the
    multi-threaded oracle."""
    from blockgen import block_image
    bw, bh = 2048, 512
    img = block_image("alpha", 2718, bw, bh)
    src = upload(img)
    for fmt in (W.DXT5, W.ETC2):
        say(f"alpha large fuzz fmt={fmt}: {bw * bh} blocks")
        got = encode_dev(gpu, fmt, src, 4 * bw, 4 * bh, 16 * bw)
        want = oracle.compress_blocks(fmt, img, 4 * bw, 4 * bh, 16 * bw)
        bad = np.nonzero((got.reshape(-1, 16) != want.reshape(-1, 16)).any(1))[0]
        assert bad.size == 0, f"{bad.size} blocks differ, first {bad[:8]}"


# ---- DXT5 and ETC2 RGBA8 from one read of the image -----------------------------------------

def test_dual_alpha_golden_digests(gpu, digests):
    im = W.ALPHA_8192
    d = digests["alpha_images"][im.name]
    src = dev_buffer(im.pitch * im.height, fill=0)
    gpu.synth_device(im.kind, im.seed, im.texture_id, im.noise, src, im.width, im.height, im.pitch)
    a, b = dev_buffer(im.blocks * 16, fill=0xCD), dev_buffer(im.blocks * 16, fill=0xCD)
    say(f"dual alpha {im.name}")
    gpu.compress_device_dual_alpha(a, b, src, im.width, im.height, im.pitch)
    import hashlib
    assert hashlib.sha256(to_host(a).tobytes()).hexdigest() == d["dxt5_sha256"]
    assert hashlib.sha256(to_host(b).tobytes()).hexdigest() == d["etc2_sha256"]


@pytest.mark.parametrize("kind,bw,bh", [("alpha", 1024, 256), ("threshold", 300, 77), ("alpha", 1 << 19, 4)])
def test_dual_alpha_vs_oracle(gpu, oracle, kind, bw, bh):
    """Random blocks; the last case has 2^19-block rows, run as stripes."""
    from blockgen import block_image
    img = block_image(kind, 91 + bh, bw, bh)
    src = upload(img)
    n = bw * bh * 16
    a, b = dev_buffer(n, fill=0xCD), dev_buffer(n, fill=0xCD)
    say(f"dual alpha {kind} {4 * bw}x{4 * bh}")
    gpu.compress_device_dual_alpha(a, b, src, 4 * bw, 4 * bh, 16 * bw)
    assert np.array_equal(to_host(a), oracle.compress_blocks(W.DXT5, img, 4 * bw, 4 * bh, 16 * bw))
    assert np.array_equal(to_host(b), oracle.compress_blocks(W.ETC2, img, 4 * bw, 4 * bh, 16 * bw))


@pytest.mark.parametrize("w,h,extra", [(48, 20, 16), (2048, 1028, 16), (4096, 2052, 0)])
def test_dual_alpha_host_matches_single_format(gpu, w, h, extra):
    stride = w * 4 + extra
    img = gpu.synth_host(W.WAVES_ALPHA, 31, 2, 9, w, h, stride, fill=0x42)
    say(f"dual alpha host {w}x{h} stride={stride}")
    a, b = gpu.compress_dual_alpha(img, w, h, stride)
    n = (w // 4) * (h // 4) * 16
    one, two = np.full(n, 0xCD, np.uint8), np.full(n, 0xCD, np.uint8)
    assert gpu.compressDXT5(one, img, w, h, stride) == 0 and gpu.compressETC2(two, img, w, h, stride) == 0
    assert np.array_equal(a, one) and np.array_equal(b, two)
