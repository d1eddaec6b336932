"""Pins the CPU oracle (oracle/goofy_oracle.c) before anything is checked against it.

Each check is chosen so that a plausible restatement mistake fails it:
golden vectors written by the reference itself, the live reference build,
closed forms the arithmetic fixes, and decode round trips with an
independent decoder.
"""
import hashlib

import numpy as np
import pytest

import goofy_amd
from goofy_amd import workloads as W


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def solid(rgb, alpha=255):
    return np.tile(np.array([*rgb, alpha], np.uint8), 16)


def le32(b, off):
    return int(b[off]) | int(b[off + 1]) << 8 | int(b[off + 2]) << 16 | int(b[off + 3]) << 24


def q5_closed(c):
    # ceil(ceil(ceil(max(c-8,0)/2)/2)/2) == ceil(max(c-8,0)/8) == (c-1)>>3 for c >= 1
    return 0 if c == 0 else (c - 1) >> 3


# ---- golden vectors (written by the reference) -------------------------------------

@pytest.mark.parametrize("fmt,key", [(W.DXT1, "dxt1"), (W.ETC1, "etc1")])
def test_golden_blocks(oracle, golden_blocks, fmt, key):
    out = oracle.encode_blocks(fmt, golden_blocks["blocks"])
    bad = np.nonzero((out != golden_blocks[key]).any(1))[0]
    assert bad.size == 0, f"{bad.size} blocks differ, first {bad[:5]}"


def test_golden_img512(oracle, golden_img512, digests):
    im = W.IMAGE_512
    img = goofy_amd.synth_host(im.kind, im.seed, im.texture_id, im.noise, im.width, im.height, im.pitch)
    assert sha(img) == digests["images"][im.name]["input_sha256"], "synthetic generator drifted"
    for fmt, key in ((W.DXT1, "dxt1"), (W.ETC1, "etc1")):
        rc, out = oracle.compress(fmt, img, im.width, im.height, im.pitch)
        assert rc == 0
        assert np.array_equal(out, golden_img512[key])


@pytest.mark.parametrize("which", ["waves-8192x8192", "stress-16384x16384-pitch65792"])
def test_golden_large_images(oracle, digests, which):
    im = {W.IMAGE_8192.name: W.IMAGE_8192, W.STRESS_16384.name: W.STRESS_16384}[which]
    img = goofy_amd.synth_host(im.kind, im.seed, im.texture_id, im.noise, im.width, im.height, im.pitch)
    d = digests["images"][im.name]
    assert sha(img) == d["input_sha256"]
    for fmt, key in ((W.DXT1, "dxt1"), (W.ETC1, "etc1")):
        out = oracle.compress_blocks(fmt, img, im.width, im.height, im.pitch)
        assert sha(out) == d[f"{key}_sha256"]


def test_golden_batch_multiple_of_4_widths(oracle, digests):
    # widths that are multiples of 4 but not 16: the reference encoded rows
    # padded to 16 pixels and the blocks were cropped; the oracle encodes
    # the blocks directly
    assert any(t["width"] % 16 for t in digests["batch"])
    for t in digests["batch"]:
        img = goofy_amd.synth_host(t["kind"], t["seed"], t["texture_id"], t["noise"], t["width"], t["height"])
        assert sha(img) == t["input_sha256"]
        out = oracle.compress_blocks(t["format"], img, t["width"], t["height"], t["width"] * 4)
        assert sha(out) == t["output_sha256"], t


# ---- live reference -------------------------------------------------------------------

@pytest.mark.parametrize("fmt", [W.DXT1, W.ETC1])
def test_fuzz_vs_reference(oracle, reference, fmt):
    rng = np.random.default_rng(1234 + fmt)
    n = 100_000
    uni = rng.integers(0, 256, (n, 64), dtype=np.uint8)
    base = rng.integers(0, 256, (n, 1))
    spread = rng.integers(0, 40, (n, 1))
    low = np.clip(base + rng.integers(-40, 41, (n, 64)) % (spread + 1), 0, 255).astype(np.uint8)
    blocks = np.concatenate([uni, low])
    a = oracle.encode_blocks(fmt, blocks)
    b = reference.encode_blocks(fmt, blocks)
    bad = np.nonzero((a != b).any(1))[0]
    assert bad.size == 0, f"{bad.size} mismatches"


@pytest.mark.parametrize("kind", ["uniform", "narrow", "ramps", "threshold"])
def test_fuzz_ten_million_blocks_vs_reference(oracle, reference, kind):
    """2,621,440 random blocks per kind (10.5M over the four kinds), both
    formats: the restatement bit-exact against the live reference."""
    from blockgen import block_image
    bw, bh = 4096, 640
    img = block_image(kind, 4321, bw, bh)
    w, h = 4 * bw, 4 * bh
    for fmt in (W.DXT1, W.ETC1):
        a = oracle.compress_blocks(fmt, img, w, h, w * 4)
        b = reference.compress_mt(fmt, img, w, h, w * 4, 8)
        bad = np.nonzero((a.reshape(-1, 8) != b.reshape(-1, 8)).any(1))[0]
        assert bad.size == 0, f"{kind} fmt={fmt}: {bad.size} blocks differ, first {bad[:8]}"


def test_reference_blocks_are_independent(reference):
    # the reference encodes 4 blocks per call; each block's bytes depend only
    # on its own pixels (what allows per-block GPU lanes and padding+crop)
    rng = np.random.default_rng(5)
    blocks = rng.integers(0, 256, (4096, 64), dtype=np.uint8)
    perm = rng.permutation(4096)
    for fmt in (W.DXT1, W.ETC1):
        a = reference.encode_blocks(fmt, blocks)
        b = reference.encode_blocks(fmt, blocks[perm])
        assert np.array_equal(a[perm], b)


def test_reference_image_equals_blocks(oracle, reference):
    im = W.IMAGE_512
    img = goofy_amd.synth_host(im.kind, im.seed, im.texture_id, im.noise, im.width, im.height, im.pitch)
    for fmt in (W.DXT1, W.ETC1):
        rc, ref_out = reference.compress(fmt, img, im.width, im.height, im.pitch)
        assert rc == 0
        assert np.array_equal(oracle.compress_blocks(fmt, img, im.width, im.height, im.pitch), ref_out)


# ---- closed forms -----------------------------------------------------------------------

def test_q5_closed_form_every_value(oracle):
    # constant block (c, c, c): c1 = q5(c) in every 5-bit field
    for c in range(256):
        out = oracle.encode_blocks(W.DXT1, solid((c, c, c)))[0]
        c0, c1 = le32(out, 0) & 0xFFFF, le32(out, 0) >> 16
        q = q5_closed(c)
        assert c1 == (q << 11) | (q << 6) | q, c
        assert c0 == c1 | 0x20


def test_constant_blocks(oracle):
    rng = np.random.default_rng(9)
    for rgb in [tuple(int(x) for x in rng.integers(0, 256, 3)) for _ in range(200)]:
        d = oracle.encode_blocks(W.DXT1, solid(rgb))[0]
        e = oracle.encode_blocks(W.ETC1, solid(rgb))[0]
        c1 = (q5_closed(rgb[0]) << 11) | (q5_closed(rgb[1]) << 6) | q5_closed(rgb[2])
        assert le32(d, 0) == (c1 | 0x20) | (c1 << 16)
        assert le32(d, 4) == 0xAAAAAAAA  # range 8, threshold 3: every pixel "near", >= mid
        assert le32(e, 4) == 0
        assert le32(e, 0) == 0x03000000 | q5_closed(rgb[0]) << 3 | q5_closed(rgb[1]) << 11 | q5_closed(rgb[2]) << 19


def test_black_and_white_blocks(oracle):
    black, white = solid((0, 0, 0)), solid((255, 255, 255))
    assert bytes(oracle.encode_blocks(W.DXT1, black)[0]) == bytes([0x20, 0, 0, 0, 0xAA, 0xAA, 0xAA, 0xAA])
    assert bytes(oracle.encode_blocks(W.ETC1, black)[0]) == bytes([0, 0, 0, 3, 0, 0, 0, 0])
    w = oracle.encode_blocks(W.DXT1, white)[0]
    assert le32(w, 0) & 0xFFFF == 0xFFFF and le32(w, 0) >> 16 == 0xFFDF and le32(w, 4) == 0xAAAAAAAA


def test_alpha_is_ignored(oracle):
    rng = np.random.default_rng(3)
    blocks = rng.integers(0, 256, (500, 16, 4), dtype=np.uint8)
    other = blocks.copy()
    other[:, :, 3] = rng.integers(0, 256, (500, 16))
    for fmt in (W.DXT1, W.ETC1):
        assert np.array_equal(oracle.encode_blocks(fmt, blocks), oracle.encode_blocks(fmt, other))


def test_dxt1_always_four_colour_mode(oracle, golden_blocks):
    out = oracle.encode_blocks(W.DXT1, golden_blocks["blocks"]).astype(np.uint32)
    c0 = out[:, 0] | out[:, 1] << 8
    c1 = out[:, 2] | out[:, 3] << 8
    assert (c0 > c1).all()


def test_etc1_control_byte_every_range(oracle, golden_blocks):
    # byte 3 of every golden ETC1S block (written by the reference) is the
    # oracle's control byte for that block's brightness range, and the golden
    # set covers every range value the encoder can produce (8..255)
    seen = {}
    for blk, out in zip(golden_blocks["blocks"], golden_blocks["etc1"]):
        r = oracle.analyse(blk)["range"]
        assert int(out[3]) == oracle.etc1_control(r), r
        seen[r] = int(out[3])
    assert set(seen) == set(range(8, 256))
    # one table index for both halves, diff and flip bits set
    assert all(v & 3 == 3 and (v >> 5) == ((v >> 2) & 7) for v in seen.values())


def test_block_analysis_brute_force(oracle):
    # bounding box, luma, mid and threshold from first principles (float-free
    # integer restatement of the README's steps) on random blocks
    rng = np.random.default_rng(11)
    avg = lambda a, b: (a + b + 1) >> 1  # noqa: E731
    for _ in range(300):
        px = rng.integers(0, 256, (16, 4))
        if rng.integers(0, 2):
            px = np.clip(px // 8 + rng.integers(100, 140), 0, 255)
        info = oracle.analyse(px.astype(np.uint8))
        mn, mx = px.min(0), px.max(0)
        ymn, ymx = avg(avg(mn[0], mn[2]), mn[1]), avg(avg(mx[0], mx[2]), mx[1])
        rng_ = max(ymx - ymn, 8)
        mid = avg(ymn, ymx)
        qt = avg(avg(rng_, 0), 0) + avg(avg(avg(rng_, 0), 0), 0)
        assert (info["ymn"], info["ymx"], info["range"], info["mid"], info["qt"]) == (ymn, ymx, rng_, mid, qt)
        y = [avg(avg(p[0], p[2]), p[1]) for p in px]
        assert info["gez"] == sum(1 << i for i in range(16) if y[i] >= mid)
        assert info["lqt"] == sum(1 << i for i in range(16) if abs(y[i] - mid) < qt)


# ---- decode round trips (independent decoder) ------------------------------------------

def test_dxt1_round_trip_classes(oracle):
    # a pixel at or above mid decodes to c0 or the 2/3 c0 point; below mid to
    # c1 or the 1/3 point; near pixels to the interpolants
    rng = np.random.default_rng(21)
    blocks = rng.integers(0, 256, (400, 64), dtype=np.uint8)
    enc = oracle.encode_blocks(W.DXT1, blocks)
    dec = oracle.decode_blocks(W.DXT1, enc)
    for i in range(400):
        info = oracle.analyse(blocks[i])
        pal = oracle.decode_blocks(W.DXT1, np.concatenate([enc[i][:4], np.array([0xE4, 0, 0, 0], np.uint8)]))[0]
        pal = pal.reshape(16, 4)[:4]  # pixels 0..3 = palette entries 0,1,2,3
        px = dec[i].reshape(16, 4)
        for p in range(16):
            ge = (info["gez"] >> p) & 1
            near = (info["lqt"] >> p) & 1
            code = (1 - ge) | (near << 1)
            assert (px[p] == pal[code]).all()


def test_round_trip_psnr(oracle):
    # decoded quality of a smooth noisy image: a wrong bit order, transposed
    # selectors or swapped endpoints drop this far below the bound
    im = W.IMAGE_512
    img = goofy_amd.synth_host(im.kind, im.seed, im.texture_id, 0, im.width, im.height)
    src = img.reshape(im.height, im.width, 4)[..., :3].astype(np.float64)
    for fmt, bound in ((W.DXT1, 25.0), (W.ETC1, 22.0)):
        rc, enc = oracle.compress(fmt, img, im.width, im.height, im.width * 4)
        dec = oracle.decode_image(fmt, enc, im.width, im.height).reshape(im.height, im.width, 4)[..., :3]
        mse = ((dec.astype(np.float64) - src) ** 2).mean()
        psnr = 10 * np.log10(255.0 ** 2 / mse)
        assert psnr > bound, (fmt, psnr)


def test_etc1_round_trip_structure(oracle):
    rng = np.random.default_rng(31)
    blocks = rng.integers(0, 256, (300, 64), dtype=np.uint8)
    enc = oracle.encode_blocks(W.ETC1, blocks)
    tables = [(2, 8), (5, 17), (9, 29), (13, 42), (18, 60), (24, 80), (33, 106), (47, 183)]
    for i in range(300):
        b = enc[i]
        assert b[3] & 3 == 3  # differential + flip
        assert (b[3] >> 5) == ((b[3] >> 2) & 7)  # one table for both halves
        assert all(x & 7 == 0 for x in b[:3])  # zero deltas
        info = oracle.analyse(blocks[i])
        dec = oracle.decode_blocks(W.ETC1, enc[i:i + 1])[0].reshape(16, 4)
        base = [((int(v) >> 3) << 3) | (int(v) >> 5) for v in b[:3]]
        a_, b_ = tables[b[3] >> 5]
        for p in range(16):
            ge = (info["gez"] >> p) & 1
            near = (info["lqt"] >> p) & 1
            mag = a_ if near else b_
            delta = mag if ge else -mag
            assert list(dec[p][:3]) == [min(255, max(0, c + delta)) for c in base]


# ---- driver semantics ----------------------------------------------------------------

def test_driver_return_codes(oracle, reference):
    img = np.zeros(64 * 64 * 4, np.uint8)
    assert oracle.compress(W.DXT1, img, 60, 64, 240)[0] == -1
    assert oracle.compress(W.DXT1, img, 64, 62, 256)[0] == -2
    assert oracle.compress(W.ETC1, img, 60, 62, 240)[0] == -1  # width is checked first
    assert oracle.compress(W.ETC1, img, 0, 64, 0)[0] == 0
    for args in ((60, 64, 240), (64, 62, 256), (60, 62, 240), (0, 64, 0)):
        assert reference.compress(W.DXT1, img, *args)[0] == oracle.compress(W.DXT1, img, *args)[0]


# ---- decoder oracle vs the reference's decoders (Src/decoder.cpp) --------------------------

def valid_etc1_blocks(rng, n):
    """Random ETC1 blocks: individual mode, or differential mode whose sums stay in 0..31."""
    out = []
    while len(out) < n:
        b = rng.integers(0, 256, 8, dtype=np.uint8)
        if b[3] & 2:
            ok = all(0 <= (int(v) >> 3) + ((int(v) & 7) - 8 * ((int(v) & 7) >= 4)) <= 31 for v in b[:3])
            if not ok:
                continue
        out.append(b)
    return np.stack(out)


def test_decoder_oracle_matches_reference_bc1(oracle, reference):
    rng = np.random.default_rng(41)
    blocks = rng.integers(0, 256, (20000, 8), dtype=np.uint8)  # both 4- and 3-colour modes
    assert np.array_equal(oracle.decode_blocks(W.DXT1, blocks), reference.decode_blocks(W.DXT1, blocks))


def test_decoder_oracle_matches_reference_etc1(oracle, reference):
    rng = np.random.default_rng(43)
    blocks = valid_etc1_blocks(rng, 20000)
    assert (blocks[:, 3] & 2).any() and not (blocks[:, 3] & 2).all()
    assert np.array_equal(oracle.decode_blocks(W.ETC1, blocks), reference.decode_blocks(W.ETC1, blocks))
    enc = oracle.encode_blocks(W.ETC1, rng.integers(0, 256, (5000, 64), dtype=np.uint8))
    assert np.array_equal(oracle.decode_blocks(W.ETC1, enc), reference.decode_blocks(W.ETC1, enc))
