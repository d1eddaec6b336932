"""TEST INFRASTRUCTURE ONLY: ctypes access to the CPU oracle.

* ``Oracle``  -- the C restatement of GoofyTC/goofy_tc.h (oracle/goofy_oracle.c),
  the rule of the DXT5 / ETC2 RGBA8 alpha blocks (no reference encoder; pinned
  through the harness's decoders by tests/test_alpha.py), and the independent
  BC1 / ETC1 / BC3 / EAC decoders.
* ``Reference`` -- the reference header compiled from its own source into
  oracle/_ref/libgoofy_ref.so (present when the reference tree was available
  at build time; the built file travels with the repository snapshot).
* ``Harness`` -- the reference benchmark harness (Src/main.cpp) compiled the
  same way into oracle/_ref/libgoofy_harness.so: its PNG loader, DDS/KTX
  writers, image decoders, getMsePsnr and the RGB565 quality floor.

Only tests/, ``__graft_entry__.smoke()`` and bench.py's CPU-baseline leg import
this module, and only as the checker.
"""
from __future__ import annotations

import ctypes
import os
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ORACLE_SO = HERE / "libgoofy_oracle.so"
REF_SO = HERE / "_ref" / "libgoofy_ref.so"
HARNESS_SO = HERE / "_ref" / "libgoofy_harness.so"

# MsePsnr field order (Src/main.cpp:367-385)
MSE_PSNR_FIELDS = ("mse_r", "mse_g", "mse_b", "mse_a", "mse_max", "mse_rgb", "mse_y",
                   "psnr_r", "psnr_g", "psnr_b", "psnr_a", "psnr_min", "psnr_rgb", "psnr_y")

DXT1, ETC1, DXT5, ETC2 = 0, 1, 2, 3


def block_bytes(fmt: int) -> int:
    """8 for DXT1 / ETC1S, 16 for DXT5 / ETC2 RGBA8 (alpha block + colour block)."""
    return 16 if fmt in (DXT5, ETC2) else 8
_u8p = ctypes.POINTER(ctypes.c_uint8)


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(_u8p)


class _BlockInfo(ctypes.Structure):
    _fields_ = [("mn", ctypes.c_uint8 * 4), ("mx", ctypes.c_uint8 * 4),
                ("ymn", ctypes.c_uint8), ("ymx", ctypes.c_uint8),
                ("range", ctypes.c_uint8), ("mid", ctypes.c_uint8), ("qt", ctypes.c_uint8),
                ("gez", ctypes.c_uint16), ("lqt", ctypes.c_uint16)]


def aligned_empty(nbytes: int, align: int = 256) -> np.ndarray:
    """uint8 buffer whose first byte is `align`-aligned."""
    raw = np.empty(nbytes + align, dtype=np.uint8)
    off = (-raw.ctypes.data) % align
    return raw[off:off + nbytes]


class Oracle:
    def __init__(self, path: Path = ORACLE_SO):
        if not path.exists():
            raise FileNotFoundError(f"{path} not built (run `make -C oracle`)")
        lib = ctypes.CDLL(str(path))
        lib.goofy_oracle_encode_block.argtypes = [ctypes.c_int, _u8p, _u8p]
        lib.goofy_oracle_analyse_block.argtypes = [_u8p, ctypes.POINTER(_BlockInfo)]
        lib.goofy_oracle_etc1_control.argtypes = [ctypes.c_uint8]
        lib.goofy_oracle_etc1_control.restype = ctypes.c_uint8
        lib.goofy_oracle_compress.argtypes = [ctypes.c_int, _u8p, _u8p, ctypes.c_uint, ctypes.c_uint, ctypes.c_uint]
        lib.goofy_oracle_compress_blocks.argtypes = [ctypes.c_int, _u8p, _u8p, ctypes.c_uint, ctypes.c_uint,
                                                     ctypes.c_size_t, ctypes.c_uint, ctypes.c_uint]
        lib.goofy_oracle_compress_mt.argtypes = [ctypes.c_int, _u8p, _u8p, ctypes.c_uint, ctypes.c_uint,
                                                 ctypes.c_size_t, ctypes.c_int]
        lib.goofy_oracle_decode_bc1.argtypes = [_u8p, _u8p]
        lib.goofy_oracle_decode_etc1.argtypes = [_u8p, _u8p]
        lib.goofy_oracle_decode_bc3.argtypes = [_u8p, _u8p]
        lib.goofy_oracle_decode_etc2.argtypes = [_u8p, _u8p]
        lib.goofy_oracle_alpha_bc3.argtypes = [_u8p, _u8p]
        lib.goofy_oracle_alpha_eac.argtypes = [_u8p, _u8p]
        lib.goofy_oracle_decode_image.argtypes = [ctypes.c_int, _u8p, _u8p, ctypes.c_uint, ctypes.c_uint]
        lib.goofy_oracle_rgb565_baseline.argtypes = [_u8p, ctypes.c_size_t, _u8p, ctypes.c_size_t, ctypes.c_uint,
                                                     ctypes.c_uint]
        self.lib = lib

    def rgb565_baseline(self, img: np.ndarray, width: int, height: int, pitch: int | None = None) -> np.ndarray:
        """The harness's quality floor; returns a tight width*height*4 image."""
        pitch = pitch or width * 4
        out = np.zeros(width * height * 4, dtype=np.uint8)
        self.lib.goofy_oracle_rgb565_baseline(_ptr(out), width * 4, _ptr(np.ascontiguousarray(img)), pitch, width,
                                              height)
        return out

    # -- blocks ---------------------------------------------------------------
    def encode_blocks(self, fmt: int, blocks: np.ndarray) -> np.ndarray:
        """blocks: (n, 64) uint8 (4 rows x 4 RGBA pixels) -> (n, 8 or 16) uint8."""
        blocks = np.ascontiguousarray(blocks, dtype=np.uint8).reshape(-1, 64)
        n = blocks.shape[0]
        # lay blocks side by side in a 4-row image and use the block driver
        img = np.ascontiguousarray(blocks.reshape(n, 4, 16).transpose(1, 0, 2)).reshape(4, n * 16)
        out = np.zeros((n, block_bytes(fmt)), dtype=np.uint8)
        if n:
            rc = self.lib.goofy_oracle_compress_blocks(fmt, _ptr(out), _ptr(img), n * 4, 4, n * 16, 0, 1)
            assert rc == 0
        return out

    def alpha_blocks(self, fmt: int, blocks: np.ndarray) -> np.ndarray:
        """The 8-byte alpha blocks alone (fmt DXT5 -> BC3, ETC2 -> EAC) of (n, 64) blocks."""
        blocks = np.ascontiguousarray(blocks, dtype=np.uint8).reshape(-1, 64)
        out = np.zeros((blocks.shape[0], 8), dtype=np.uint8)
        fn = self.lib.goofy_oracle_alpha_bc3 if fmt == DXT5 else self.lib.goofy_oracle_alpha_eac
        for i in range(blocks.shape[0]):
            fn(_ptr(blocks[i]), _ptr(out[i]))
        return out

    def analyse(self, block: np.ndarray) -> dict:
        b = np.ascontiguousarray(block, dtype=np.uint8).reshape(64)
        info = _BlockInfo()
        self.lib.goofy_oracle_analyse_block(_ptr(b), ctypes.byref(info))
        return {"mn": bytes(info.mn), "mx": bytes(info.mx), "ymn": info.ymn, "ymx": info.ymx,
                "range": info.range, "mid": info.mid, "qt": info.qt, "gez": info.gez, "lqt": info.lqt}

    def etc1_control(self, rng: int) -> int:
        return int(self.lib.goofy_oracle_etc1_control(rng))

    # -- images -----------------------------------------------------------------
    def compress(self, fmt: int, img: np.ndarray, width: int, height: int, stride: int) -> tuple[int, np.ndarray]:
        """Drop-in semantics (-1 for width % 16, -2 for height % 4)."""
        n = (width // 4) * (height // 4) * block_bytes(fmt)
        out = np.zeros(max(1, n), dtype=np.uint8)
        rc = self.lib.goofy_oracle_compress(fmt, _ptr(out), _ptr(img), width, height, stride)
        return rc, out[:n]

    def compress_blocks(self, fmt: int, img: np.ndarray, width: int, height: int, stride: int,
                        threads: int = 0) -> np.ndarray:
        """Per-block semantics for width % 4 == 0 (multi-threaded)."""
        n = (width // 4) * (height // 4) * block_bytes(fmt)
        out = np.zeros(max(1, n), dtype=np.uint8)
        threads = threads or min(16, os.cpu_count() or 1)
        rc = self.lib.goofy_oracle_compress_mt(fmt, _ptr(out), _ptr(img), width, height, stride, threads)
        if rc != 0:
            raise ValueError(f"oracle rejected {width}x{height}: {rc}")
        return out[:n]

    def decode_blocks(self, fmt: int, blocks: np.ndarray) -> np.ndarray:
        blocks = np.ascontiguousarray(blocks, dtype=np.uint8).reshape(-1, block_bytes(fmt))
        out = np.zeros((blocks.shape[0], 64), dtype=np.uint8)
        fn = {DXT1: self.lib.goofy_oracle_decode_bc1, ETC1: self.lib.goofy_oracle_decode_etc1,
              DXT5: self.lib.goofy_oracle_decode_bc3, ETC2: self.lib.goofy_oracle_decode_etc2}[fmt]
        for i in range(blocks.shape[0]):
            fn(_ptr(blocks[i]), _ptr(out[i]))
        return out

    def decode_image(self, fmt: int, blocks: np.ndarray, width: int, height: int) -> np.ndarray:
        out = np.zeros(width * height * 4, dtype=np.uint8)
        blocks = np.ascontiguousarray(blocks, dtype=np.uint8)
        rc = self.lib.goofy_oracle_decode_image(fmt, _ptr(out), _ptr(blocks), width, height)
        assert rc == 0
        return out


class Reference:
    """The reference encoder (goofy_tc.h compiled from its own source)."""

    def __init__(self, path: Path = REF_SO):
        if not path.exists():
            raise FileNotFoundError(f"{path} not built (reference tree absent at build time)")
        lib = ctypes.CDLL(str(path))
        lib.goofy_ref_compress.argtypes = [ctypes.c_int, _u8p, _u8p, ctypes.c_uint, ctypes.c_uint, ctypes.c_uint]
        lib.goofy_ref_compress_mt.argtypes = [ctypes.c_int, _u8p, _u8p, ctypes.c_uint, ctypes.c_uint,
                                              ctypes.c_uint, ctypes.c_int]
        lib.goofy_ref_encode_blocks.argtypes = [ctypes.c_int, _u8p, ctypes.c_size_t, _u8p]
        lib.goofy_ref_decode_block.argtypes = [ctypes.c_int, _u8p, _u8p]
        self.lib = lib

    def decode_blocks(self, fmt: int, blocks: np.ndarray) -> np.ndarray:
        """Reference decoders (Src/decoder.cpp) on (n, 8) blocks -> (n, 64) RGBA8."""
        blocks = np.ascontiguousarray(blocks, dtype=np.uint8).reshape(-1, 8)
        out = np.zeros((blocks.shape[0], 64), dtype=np.uint8)
        for i in range(blocks.shape[0]):
            self.lib.goofy_ref_decode_block(fmt, _ptr(blocks[i]), _ptr(out[i]))
        return out

    def encode_blocks(self, fmt: int, blocks: np.ndarray) -> np.ndarray:
        blocks = np.ascontiguousarray(blocks, dtype=np.uint8).reshape(-1, 64)
        out = np.zeros((blocks.shape[0], 8), dtype=np.uint8)
        if blocks.shape[0]:
            rc = self.lib.goofy_ref_encode_blocks(fmt, _ptr(blocks), blocks.shape[0], _ptr(out))
            assert rc == 0
        return out

    def compress(self, fmt: int, img: np.ndarray, width: int, height: int, stride: int) -> tuple[int, np.ndarray]:
        out = np.zeros(max(1, (width // 4) * (height // 4) * 8), dtype=np.uint8)
        rc = self.lib.goofy_ref_compress(fmt, _ptr(out), _ptr(img), width, height, stride)
        return rc, out[: (width // 4) * (height // 4) * 8]

    def compress_mt(self, fmt: int, img: np.ndarray, width: int, height: int, stride: int, threads: int,
                    out: np.ndarray | None = None) -> np.ndarray:
        n = (width // 4) * (height // 4) * 8
        if out is None:
            out = np.zeros(n, dtype=np.uint8)
        rc = self.lib.goofy_ref_compress_mt(fmt, _ptr(out), _ptr(img), width, height, stride, threads)
        if rc != 0:
            raise ValueError(f"reference rejected input: {rc}")
        return out


class Harness:
    """The reference benchmark harness (Src/main.cpp) compiled from its own
    source; see oracle/harness_shim.cpp."""

    def __init__(self, path: Path = HARNESS_SO):
        if not path.exists():
            raise FileNotFoundError(f"{path} not built (reference tree absent at build time)")
        lib = ctypes.CDLL(str(path))
        uintp = ctypes.POINTER(ctypes.c_uint)
        lib.goofy_harness_load_png.argtypes = [ctypes.c_char_p, uintp, uintp]
        lib.goofy_harness_load_png.restype = ctypes.c_void_p
        lib.goofy_harness_free.argtypes = [ctypes.c_void_p]
        lib.goofy_harness_mse_psnr.argtypes = [_u8p, _u8p, ctypes.c_uint, ctypes.c_uint,
                                               ctypes.POINTER(ctypes.c_double)]
        lib.goofy_harness_rgb565_baseline.argtypes = [_u8p, ctypes.c_uint, ctypes.c_uint, _u8p]
        lib.goofy_harness_decompress.argtypes = [ctypes.c_int, _u8p, ctypes.c_uint, ctypes.c_uint, _u8p]
        lib.goofy_harness_save.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_char_p, _u8p, ctypes.c_size_t,
                                           ctypes.c_uint, ctypes.c_uint]
        lib.goofy_harness_run_test.argtypes = [ctypes.c_int, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_void_p,
                                               ctypes.c_uint, _u8p, ctypes.c_uint, ctypes.c_uint, ctypes.c_uint, _u8p,
                                               ctypes.c_size_t, ctypes.POINTER(ctypes.c_double),
                                               ctypes.POINTER(ctypes.c_double)]
        lib.goofy_harness_goofy_encoder.argtypes = [ctypes.c_int]
        lib.goofy_harness_goofy_encoder.restype = ctypes.c_void_p
        self.lib = lib

    def load_png(self, path) -> tuple[np.ndarray, int, int] | None:
        """loadPngAsRgba8: (RGBA8 copy, width, height), or None when refused."""
        w, h = ctypes.c_uint(0), ctypes.c_uint(0)
        p = self.lib.goofy_harness_load_png(str(path).encode(), ctypes.byref(w), ctypes.byref(h))
        if not p:
            return None
        try:
            n = w.value * h.value * 4
            img = np.ctypeslib.as_array(ctypes.cast(p, _u8p), shape=(n,)).copy()
        finally:
            self.lib.goofy_harness_free(p)
        return img, w.value, h.value

    def mse_psnr(self, a: np.ndarray, b: np.ndarray, width: int, height: int) -> dict:
        """getMsePsnr over two tight RGBA8 images."""
        out = (ctypes.c_double * 14)()
        self.lib.goofy_harness_mse_psnr(_ptr(np.ascontiguousarray(a)), _ptr(np.ascontiguousarray(b)), width, height,
                                        out)
        return dict(zip(MSE_PSNR_FIELDS, list(out)))

    def rgb565_baseline(self, img: np.ndarray, width: int, height: int) -> np.ndarray:
        out = np.zeros(width * height * 4, dtype=np.uint8)
        self.lib.goofy_harness_rgb565_baseline(_ptr(np.ascontiguousarray(img)), width, height, _ptr(out))
        return out

    def decompress(self, fmt: int, blocks: np.ndarray, width: int, height: int) -> np.ndarray:
        out = np.zeros(width * height * 4, dtype=np.uint8)
        self.lib.goofy_harness_decompress(fmt, _ptr(np.ascontiguousarray(blocks)), width, height, _ptr(out))
        return out

    def goofy_encoder(self, fmt: int) -> int:
        """Address of goofy::compressDXT1 (fmt 0) / compressETC1 (fmt 1) as the
        harness compiles them: a CompressFunc_t for run_test."""
        return int(self.lib.goofy_harness_goofy_encoder(fmt))

    def run_test(self, fmt: int, encoder: str, image: str, func: int, iterations: int, img: np.ndarray,
                 width: int, height: int, stride: int) -> tuple[np.ndarray, float, dict]:
        """runTestDXT1 / runTestETC1 (Src/main.cpp:646-743) around the
        CompressFunc_t at address `func`; writes the harness's files under
        ./test-results/.  Returns (blocks, best microseconds, getMsePsnr)."""
        src = np.ascontiguousarray(img, dtype=np.uint8)
        dst = np.zeros((width // 4) * (height // 4) * 8, dtype=np.uint8)
        t = ctypes.c_double(0)
        q = (ctypes.c_double * 14)()
        self.lib.goofy_harness_run_test(fmt, encoder.encode(), image.encode(), func, iterations, _ptr(src), width,
                                        height, stride, _ptr(dst), dst.size, ctypes.byref(t), q)
        return dst, t.value, dict(zip(MSE_PSNR_FIELDS, list(q)))

    def save(self, container: int, fmt: int, path, data: np.ndarray, width: int, height: int) -> None:
        """saveDds (container 0) / saveKtx (container 1) to `path`."""
        data = np.ascontiguousarray(data, dtype=np.uint8)
        self.lib.goofy_harness_save(container, fmt, str(path).encode(), _ptr(data), data.size, width, height)


def have_reference() -> bool:
    return REF_SO.exists()


def have_harness() -> bool:
    return HARNESS_SO.exists()
