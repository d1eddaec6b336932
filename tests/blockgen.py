"""Random 4x4 RGBA8 blocks of four kinds, laid out as an image, for fuzzing
the encoder against the reference.

* uniform   -- uniform random bytes (ranges near the top of 8..255);
* narrow    -- small spreads around random colours (every range 8..48);
* ramps     -- per-channel gradients with a little noise;
* threshold -- grey blocks whose lumas sit on eighths of a random range, so
               pixels land on, just inside and just outside the >= mid and
               near/far boundaries of every range;
* alpha     -- random colour with alpha of six kinds per block: uniform bytes,
               uniform within a random [lo, hi], 0/255 cutout, constant, ramp
               lo..hi, and lo with a little noise (the DXT5 / ETC2 inputs).

Blocks are generated in chunks with small integer types, so that images of
millions of blocks stay within a few hundred MiB.
"""
from __future__ import annotations

import numpy as np

KINDS = ("uniform", "narrow", "ramps", "threshold", "alpha")
_CHUNK = 1 << 18  # blocks per generation step


def _blocks(kind: str, rng: np.random.Generator, n: int) -> np.ndarray:
    """(n, 16, 4) uint8: n blocks, 16 pixels each in row-major order."""
    if kind == "uniform":
        return rng.integers(0, 256, (n, 16, 4), dtype=np.uint8)
    if kind == "narrow":
        base = rng.integers(0, 256, (n, 1, 4), dtype=np.int16)
        spread = rng.integers(0, 25, (n, 1, 1), dtype=np.int16)
        px = base + rng.integers(-24, 25, (n, 16, 4), dtype=np.int16) * spread // 24
    elif kind == "ramps":
        t = np.arange(16, dtype=np.int32)[None, :, None]
        a = rng.integers(0, 256, (n, 1, 4), dtype=np.int32)
        b = rng.integers(0, 256, (n, 1, 4), dtype=np.int32)
        px = a + (b - a) * t // 15 + rng.integers(-2, 3, (n, 16, 4), dtype=np.int32)
    elif kind == "threshold":
        lo = rng.integers(0, 256, (n, 1), dtype=np.int32)
        r = rng.integers(0, 256, (n, 1), dtype=np.int32) * (255 - lo) // 255  # lo + r <= 255
        g = lo + r * rng.integers(0, 9, (n, 16), dtype=np.int32) // 8
        px = np.repeat(g[..., None], 4, axis=2)
        px[..., 3] = rng.integers(0, 256, (n, 16), dtype=np.int32)  # alpha never reaches the output
    elif kind == "alpha":
        px = rng.integers(0, 256, (n, 16, 4), dtype=np.int32)
        k = rng.integers(0, 6, (n, 1))
        lo = rng.integers(0, 256, (n, 1), dtype=np.int32)
        hi = rng.integers(0, 256, (n, 1), dtype=np.int32)
        lo, hi = np.minimum(lo, hi), np.maximum(lo, hi)
        a = px[:, :, 3]
        a = np.where(k == 1, lo + rng.integers(0, 1 << 16, (n, 16)) % (hi - lo + 1), a)
        a = np.where(k == 2, np.where(rng.random((n, 16)) < 0.5, 0, 255), a)
        a = np.where(k == 3, lo, a)
        a = np.where(k == 4, lo + (hi - lo) * np.arange(16)[None, :] // 15, a)
        a = np.where(k == 5, lo + rng.integers(-6, 7, (n, 16)), a)
        px[:, :, 3] = a
    else:
        raise ValueError(kind)
    return np.clip(px, 0, 255).astype(np.uint8)


def random_blocks(kind: str, seed: int, n: int) -> np.ndarray:
    """(n, 64) uint8 blocks of one kind, each 16 RGBA8 pixels in row-major order."""
    return _blocks(kind, np.random.default_rng(seed), n).reshape(n, 64)


def block_image(kind: str, seed: int, bw: int, bh: int) -> np.ndarray:
    """bw x bh independent random blocks of one kind as a 64-byte-aligned flat
    RGBA8 image of width 4*bw and pitch 16*bw bytes."""
    from oracle.oracle import aligned_empty
    rng = np.random.default_rng(seed)
    img = aligned_empty(bw * bh * 64)
    view = img.reshape(bh, 4, bw, 16)  # block row, pixel row, block, 4 pixels x 4 bytes
    rows_per_chunk = max(1, _CHUNK // bw)
    for r0 in range(0, bh, rows_per_chunk):
        r1 = min(bh, r0 + rows_per_chunk)
        px = _blocks(kind, rng, (r1 - r0) * bw).reshape(r1 - r0, bw, 4, 16)  # block row, block, y, x*ch
        view[r0:r1] = px.transpose(0, 2, 1, 3)
    return img
