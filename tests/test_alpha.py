"""DXT5 (BC3) and ETC2 RGBA8 (EAC alpha) blocks on the CPU: the oracle's
alpha encoders and decoders, pinned through the reference's own decoders.

The reference has no alpha encoder (its README names one, "alpha directly
instead of brightness", as the next step), so nothing can be bit-exact
against a reference encoder.  What pins the alpha blocks instead:

* the colour half of every 16-byte block equals the reference's DXT1 / ETC1S
  block of the same pixels (the reference encoder, compiled from source);
* the oracle's BC3 / EAC decoders equal the harness's decompressDXT5 /
  decompressETC2 (Src/main.cpp:579-631 over Src/decoder.cpp) on random blocks;
* through the harness's decoders, every pixel of every encoded block decodes
  to the palette level nearest to it (ties to the higher level), where the
  palette is read back from the same decoder by forcing all 16 selectors to
  each value in turn -- for every (min, max) pair of alpha values and every
  alpha in between;
* the palette covers [min, max]; constant blocks, and blocks of only 0 and
  255, come back exact.
"""
import numpy as np
import pytest

from blockgen import random_blocks
from oracle.oracle import DXT1, DXT5, ETC1, ETC2, block_bytes

ALPHA_FORMATS = [(DXT5, DXT1), (ETC2, ETC1)]


def decode_ref(harness, fmt: int, blocks: np.ndarray) -> np.ndarray:
    """Harness decoder on (n, 16) blocks -> (n, 16) decoded alpha (row-major)."""
    n = blocks.shape[0]
    img = harness.decompress(fmt, np.ascontiguousarray(blocks).reshape(-1), 4 * n, 4).reshape(4, n, 4, 4)
    return img[..., 3].transpose(1, 0, 2).reshape(n, 16)


def forced(blocks: np.ndarray, fmt: int, sel: int) -> np.ndarray:
    """Copies of the blocks with all 16 alpha selectors set to `sel`."""
    out = blocks.copy()
    bits = 0
    for p in range(16):
        bits |= sel << (3 * p)  # same pattern for every pixel: bit order does not matter
    out[:, 2:8] = np.frombuffer(bits.to_bytes(6, "little" if fmt == DXT5 else "big"), np.uint8)
    return out


def palettes(harness, fmt: int, blocks: np.ndarray) -> np.ndarray:
    """(n, 8) alpha palette of each block, as the harness's decoder forms it."""
    return np.stack([decode_ref(harness, fmt, forced(blocks, fmt, s))[:, 0] for s in range(8)], axis=1)


# ---- the oracle's decoders against the harness's -------------------------------------------

@pytest.mark.parametrize("fmt,cfmt", ALPHA_FORMATS)
def test_oracle_decoders_match_harness(oracle, harness, fmt, cfmt):
    rng = np.random.default_rng(5 + fmt)
    n = 4096
    colour = oracle.encode_blocks(cfmt, rng.integers(0, 256, (n, 64), dtype=np.uint8))
    alpha = rng.integers(0, 256, (n, 8), dtype=np.uint8)  # every base, multiplier, table and selector
    if fmt == DXT5:
        alpha[: n // 4, 1] = alpha[: n // 4, 0]  # a0 == a1 (6-level mode) ...
        alpha[n // 4: n // 2, :2].sort(axis=1)   # ... and a0 < a1
    blocks = np.concatenate([alpha, colour], axis=1)
    ours = oracle.decode_image(fmt, blocks.reshape(-1), 4 * n, 4)
    theirs = harness.decompress(fmt, blocks.reshape(-1), 4 * n, 4)
    assert np.array_equal(ours, theirs)


# ---- the 16-byte blocks ------------------------------------------------------------------------

@pytest.mark.parametrize("fmt,cfmt", ALPHA_FORMATS)
def test_colour_half_is_the_reference_block(oracle, reference, fmt, cfmt):
    blocks = random_blocks("alpha", 11 + fmt, 20000)
    out = oracle.encode_blocks(fmt, blocks)
    assert out.shape == (20000, block_bytes(fmt))
    assert np.array_equal(out[:, 8:], reference.encode_blocks(cfmt, blocks))
    assert np.array_equal(out[:512, :8], oracle.alpha_blocks(fmt, blocks[:512]))


def check_nearest(harness, fmt, blocks64, enc):
    """Every pixel decodes to its nearest palette level, ties to the higher one."""
    a = blocks64.reshape(-1, 16, 4)[:, :, 3].astype(np.int32)
    dec = decode_ref(harness, fmt, enc).astype(np.int32)
    pal = palettes(harness, fmt, enc).astype(np.int32)
    dist = np.abs(pal[:, None, :] - a[:, :, None])                      # (n, 16, 8)
    best = dist.min(axis=2)
    assert np.array_equal(np.abs(dec - a), best), "a pixel did not take its nearest level"
    top = np.where(dist == best[..., None], pal[:, None, :], -1).max(axis=2)
    assert np.array_equal(dec, top), "a tie did not go to the higher level"
    return a, dec, pal


@pytest.mark.parametrize("fmt", [DXT5, ETC2])
def test_alpha_nearest_level_random(oracle, harness, fmt):
    blocks = random_blocks("alpha", 23 + fmt, 30000)
    enc = oracle.encode_blocks(fmt, blocks)
    a, dec, pal = check_nearest(harness, fmt, blocks, enc)
    assert (pal.min(1) <= a.min(1)).all() and (pal.max(1) >= a.max(1)).all()


def all_pairs_blocks() -> np.ndarray:
    """For every 0 <= lo <= hi <= 255, blocks holding lo, hi and every value in
    between (14 free pixels per block); random colour."""
    rng = np.random.default_rng(3)
    rows = []
    for lo in range(256):
        for hi in range(lo, 256):
            vals = np.arange(lo, hi + 1)
            nb = max(1, -(-vals.size // 14))
            fill = np.resize(vals, nb * 14).reshape(nb, 14)
            rows.append(np.concatenate([np.full((nb, 1), lo), np.full((nb, 1), hi), fill], axis=1))
    a = np.concatenate(rows)
    perm = rng.permuted(np.tile(np.arange(16), (a.shape[0], 1)), axis=1)  # lo/hi anywhere in the block
    a = np.take_along_axis(a, perm, axis=1)
    px = rng.integers(0, 256, (a.shape[0], 16, 4), dtype=np.int32)
    px[:, :, 3] = a
    return px.astype(np.uint8).reshape(-1, 64)


@pytest.fixture(scope="module")
def pair_blocks():
    return all_pairs_blocks()


@pytest.mark.parametrize("fmt", [DXT5, ETC2])
def test_alpha_nearest_level_every_range(oracle, harness, pair_blocks, fmt):
    enc = oracle.encode_blocks(fmt, pair_blocks)
    a, dec, pal = check_nearest(harness, fmt, pair_blocks, enc)
    lo, hi = a.min(1), a.max(1)
    assert (pal.min(1) <= lo).all() and (pal.max(1) >= hi).all()
    err = np.abs(dec - a).max(1)
    if fmt == DXT5:
        assert (enc[:, 0] == hi).all() and (enc[:, 1] == lo).all()
        assert (err <= (hi - lo) // 14 + 1).all()  # half a seventh of the range, plus the truncation
    else:
        m = enc[:, 1] >> 4
        assert ((enc[:, 1] & 15) == 9).all()
        assert np.array_equal(m, np.maximum(1, -(-(hi.astype(int) - lo) // 19)))
        assert (err <= (3 * m + 1) // 2).all()  # half the widest gap of table 9 (3 steps)


@pytest.mark.parametrize("fmt", [DXT5, ETC2])
def test_alpha_exact_cases(oracle, harness, fmt):
    rng = np.random.default_rng(7)
    const = np.repeat(np.arange(256), 16).reshape(256, 16)
    binary = np.where(rng.random((4096, 16)) < 0.5, 0, 255)
    binary[:2] = [[0] * 16, [255] * 16]
    a = np.concatenate([const, binary])
    px = rng.integers(0, 256, (a.shape[0], 16, 4), dtype=np.int32)
    px[:, :, 3] = a
    blocks = px.astype(np.uint8).reshape(-1, 64)
    dec = decode_ref(harness, fmt, oracle.encode_blocks(fmt, blocks))
    assert np.array_equal(dec, a)


@pytest.mark.parametrize("fmt", [DXT5, ETC2])
def test_alpha_drop_in_semantics(oracle, fmt):
    rng = np.random.default_rng(fmt)
    w, h, stride = 48, 20, 48 * 4 + 16
    img = rng.integers(0, 256, stride * h, dtype=np.uint8)
    rc, out = oracle.compress(fmt, img, w, h, stride)
    assert rc == 0 and out.size == (w // 4) * (h // 4) * 16
    blocks = np.stack([img[(4 * by + r) * stride + 16 * bx:(4 * by + r) * stride + 16 * bx + 16]
                       for by in range(h // 4) for bx in range(w // 4) for r in range(4)]).reshape(-1, 64)
    assert np.array_equal(out.reshape(-1, 16), oracle.encode_blocks(fmt, blocks))
    assert oracle.compress(fmt, img, 44, 20, stride)[0] == -1
    assert oracle.compress(fmt, img, 48, 18, stride)[0] == -2
