"""Inputs and outputs either side of the encoder.

* ``load_png_rgba8`` -- PNG -> 64-byte-aligned RGBA8 with the rules of the
  reference harness's loader (loadPngAsRgba8, Src/main.cpp:258-343): any PNG
  colour type becomes 8-bit RGBA, alpha is forced to 255, and widths that
  are not a multiple of 16 / heights not a multiple of 4 are refused.
* ``container_bytes`` / ``write_container`` -- DDS (DXT1, DXT5) and KTX1 (any
  format) files with the harness's headers (saveDds / saveKtx,
  Src/main.cpp:154-220), written by libgoofy_hip.so.
"""
from __future__ import annotations

import ctypes
from pathlib import Path

import numpy as np

from . import _check, _lib

DDS, KTX = 0, 1

_lib.goofy_container_bytes.restype = ctypes.c_size_t
_lib.goofy_container_bytes.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_uint, ctypes.c_uint]
_lib.goofy_container_write.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_uint, ctypes.c_uint, ctypes.c_void_p,
                                       ctypes.c_size_t, ctypes.c_void_p, ctypes.c_size_t,
                                       ctypes.POINTER(ctypes.c_size_t)]
_lib.goofy_container_write_file.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.c_uint,
                                            ctypes.c_uint, ctypes.c_void_p, ctypes.c_size_t]


class ImageRejected(ValueError):
    pass


def load_png_rgba8(path: str | Path) -> tuple[np.ndarray, int, int]:
    """(flat uint8 RGBA8 buffer, 64-byte aligned; width; height)."""
    from PIL import Image  # PNG decoding only; the encoder never sees PIL objects

    with Image.open(path) as im:
        if im.mode.startswith("I"):
            # 16-bit greyscale: the harness's decoder (lodepng, to 8-bit RGBA)
            # keeps the high byte of each sample; PIL's RGBA conversion would clip
            g = (np.asarray(im).astype(np.uint32) >> 8).astype(np.uint8)
            rgba = np.stack([g, g, g, np.full_like(g, 255)], axis=-1)
        else:
            rgba = np.asarray(im.convert("RGBA"), dtype=np.uint8)
    h, w = rgba.shape[:2]
    if w % 16 != 0:
        raise ImageRejected(f"{path}: width {w} is not a multiple of 16")
    if h % 4 != 0:
        raise ImageRejected(f"{path}: height {h} is not a multiple of 4")
    raw = np.empty(w * h * 4 + 64, dtype=np.uint8)
    off = (-raw.ctypes.data) % 64
    out = raw[off:off + w * h * 4]
    out[:] = rgba.reshape(-1)
    out[3::4] = 255
    return out, w, h


def container_bytes(kind: int, fmt: int, width: int, height: int, blocks: np.ndarray) -> bytes:
    blocks = np.ascontiguousarray(blocks, dtype=np.uint8)
    cap = int(_lib.goofy_container_bytes(kind, fmt, width, height))
    out = np.empty(cap, dtype=np.uint8)
    n = ctypes.c_size_t()
    _check(_lib.goofy_container_write(kind, fmt, width, height, blocks.ctypes.data, blocks.nbytes, out.ctypes.data,
                                      cap, ctypes.byref(n)))
    return out[: n.value].tobytes()


def write_container(path: str | Path, kind: int, fmt: int, width: int, height: int, blocks: np.ndarray) -> None:
    blocks = np.ascontiguousarray(blocks, dtype=np.uint8)
    _check(_lib.goofy_container_write_file(str(path).encode(), kind, fmt, width, height, blocks.ctypes.data,
                                           blocks.nbytes))
