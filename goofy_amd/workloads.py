"""Concrete, seeded workloads of the encoder (sizes, seeds, formats).

* ``IMAGE_512``   -- one 512x512 RGBA8 texture, stride 2048 B: waves + noise 12
                     (the reference's own CPU-runnable case).
* ``IMAGE_8192``  -- one 8192x8192 RGBA8 texture, stride 32768 B: waves + noise 12.
* ``batch_textures()`` -- 4096 textures, width and height drawn independently,
                     log-uniform over multiples of 4 in [64, 2048], 50/50
                     DXT1/ETC1S, per-texture noise amplitude uniform in [0, 40].
* ``serve_pool()`` -- request textures, log-uniform over multiples of 4 in
                     [64, 4096] (width and height independent), 50/50 formats.
* ``STRESS_16384`` -- one 16384x16384 RGBA8 texture, pitch 65792 B (a multiple
                     of 256 above 4*width), built from 64x64 stress regions.
* ``ALPHA_512`` / ``ALPHA_8192`` -- waves + noise 12 with an alpha channel
                     (64x64 regions opaque, binary cutout or smooth), the
                     inputs of the DXT5 / ETC2 RGBA8 encoders.

Texture content comes from the library's integer generator (goofy_synth_*),
so host and device produce identical bytes for the same parameters.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

WAVES, STRESS, WAVES_ALPHA = 0, 1, 2
DXT1, ETC1, DXT5, ETC2 = 0, 1, 2, 3


@dataclass(frozen=True)
class Image:
    name: str
    kind: int
    seed: int
    texture_id: int
    noise: int
    width: int
    height: int
    pitch: int

    @property
    def pixels(self) -> int:
        return self.width * self.height

    @property
    def blocks(self) -> int:
        return (self.width // 4) * (self.height // 4)


IMAGE_512 = Image("waves-512x512", WAVES, 0x601F0001, 0, 12, 512, 512, 2048)
IMAGE_8192 = Image("waves-8192x8192", WAVES, 0x601F0002, 0, 12, 8192, 8192, 32768)
STRESS_16384 = Image("stress-16384x16384-pitch65792", STRESS, 0x601F0005, 0, 0, 16384, 16384, 65792)
ALPHA_512 = Image("waves-alpha-512x512", WAVES_ALPHA, 0x601F0006, 0, 12, 512, 512, 2048)
ALPHA_8192 = Image("waves-alpha-8192x8192", WAVES_ALPHA, 0x601F0007, 0, 12, 8192, 8192, 32768)

BATCH_SEED = 0x601F0003
SERVE_SEED = 0x601F0004


def _log_uniform_mult4(rng: np.random.Generator, lo: int, hi: int, n: int) -> np.ndarray:
    v = np.exp(rng.uniform(np.log(lo), np.log(hi), size=n))
    return np.clip(np.rint(v / 4.0).astype(np.int64) * 4, lo, hi)


def batch_textures(n: int = 4096, seed: int = BATCH_SEED, lo: int = 64, hi: int = 2048) -> list[Image | dict]:
    """Texture list of the batched workload: dicts with width, height, format, noise, seed, texture_id."""
    rng = np.random.default_rng(seed)
    w = _log_uniform_mult4(rng, lo, hi, n)
    h = _log_uniform_mult4(rng, lo, hi, n)
    fmt = rng.integers(0, 2, size=n)
    noise = rng.integers(0, 41, size=n)
    return [dict(width=int(w[i]), height=int(h[i]), format=int(fmt[i]), noise=int(noise[i]),
                 kind=WAVES, seed=seed, texture_id=i) for i in range(n)]


def serve_pool(n: int = 512, seed: int = SERVE_SEED, lo: int = 64, hi: int = 4096) -> list[dict]:
    """Distinct request textures of the serving workload (requests draw from this pool)."""
    rng = np.random.default_rng(seed)
    w = _log_uniform_mult4(rng, lo, hi, n)
    h = _log_uniform_mult4(rng, lo, hi, n)
    fmt = rng.integers(0, 2, size=n)
    noise = rng.integers(0, 41, size=n)
    return [dict(width=int(w[i]), height=int(h[i]), format=int(fmt[i]), noise=int(noise[i]),
                 kind=WAVES, seed=seed, texture_id=i) for i in range(n)]


def algorithmic_bytes(width: int, height: int) -> int:
    """Bytes an encode must move: 64 B of pixels in + 8 B out per 4x4 block."""
    return (width // 4) * (height // 4) * 72
