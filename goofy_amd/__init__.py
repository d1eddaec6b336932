"""goofy_amd -- Python host mirror of the Goofy encoder API on MI355X.

Everything here is a thin ctypes layer over the C-ABI of ``libgoofy_hip.so``
(declared in ``include/goofy_hip.h``); all encoding runs in the HIP kernels of
that library.  There is no CPU fallback: importing this package without the
built library raises ImportError.

Reference interfaces mirrored:
  * ``compressDXT1`` / ``compressETC1`` -- ``goofy::compressDXT1/ETC1``
    (GoofyTC/goofy_tc.h:10-13, :1497-1557): same argument meaning, same
    return codes (-1 width % 16, -2 height % 4), host buffers.
  * ``CompressFunc`` -- the harness plugin type ``CompressFunc_t``
    (Src/main.cpp:644): the two C entry points with exactly that signature.
"""
from __future__ import annotations

import ctypes
from pathlib import Path
from typing import Iterable, Sequence

import numpy as np

_LIB_PATH = Path(__file__).resolve().parent / "libgoofy_hip.so"

DXT1 = 0
ETC1 = 1
DXT5 = 2   # BC3: alpha block + the DXT1 block (16 bytes)
ETC2 = 3   # ETC2 RGBA8: EAC alpha block + the ETC1S block (16 bytes)
FORMATS = {"dxt1": DXT1, "etc1": ETC1, "etc1s": ETC1, "dxt5": DXT5, "etc2": ETC2}


def block_bytes(fmt: int) -> int:
    """Bytes per 4x4 block: 8 for DXT1 / ETC1S, 16 for DXT5 / ETC2 RGBA8."""
    return 16 if fmt in (DXT5, ETC2) else 8

OK = 0
ERR_WIDTH = -1
ERR_HEIGHT = -2
ERR_HIP = -3
ERR_INVALID = -4
ERR_OUT_OF_MEMORY = -5
ERR_NO_DEVICE = -6
ERR_TOO_LARGE = -7
ERR_BUSY = -8
ERR_CANCELLED = -9
ERR_DEADLINE = -10


class GoofyError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"goofy status {status}: {message}")
        self.status = status


if not _LIB_PATH.exists():
    raise ImportError(f"{_LIB_PATH} is not built: run `python goofy_amd/build.py` "
                      "(the encoder has no CPU fallback)")

_lib = ctypes.CDLL(str(_LIB_PATH))
_u8p = ctypes.POINTER(ctypes.c_uint8)
_vp = ctypes.c_void_p
_uint = ctypes.c_uint
_sz = ctypes.c_size_t

CompressFunc = ctypes.CFUNCTYPE(ctypes.c_int, _u8p, _u8p, _uint, _uint, _uint)


class LaunchConfig(ctypes.Structure):
    _fields_ = [("blocks_per_lane", ctypes.c_int), ("wg_waves", ctypes.c_int),
                ("waves_per_cu", ctypes.c_int), ("nontemporal", ctypes.c_int), ("tile_order", ctypes.c_int)]

    def as_dict(self) -> dict:
        return {f: getattr(self, f) for f, _ in self._fields_}


class TextureDesc(ctypes.Structure):
    _fields_ = [("src", _vp), ("dst", _vp), ("width", _uint), ("height", _uint),
                ("pitch", _sz), ("format", ctypes.c_int)]


def _sig(name, restype, *argtypes):
    fn = getattr(_lib, name)
    fn.restype = restype
    fn.argtypes = list(argtypes)
    return fn


_sig("goofy_hip_last_error", ctypes.c_char_p)
_sig("goofy_hip_status_string", ctypes.c_char_p, ctypes.c_int)
_sig("goofy_hip_version", ctypes.c_char_p)
_sig("goofy_hip_device_count", ctypes.c_int)
_sig("goofy_hip_compress", ctypes.c_int, ctypes.c_int, _u8p, _u8p, _uint, _uint, _uint)
_sig("goofy_hip_compress_dxt1", ctypes.c_int, _u8p, _u8p, _uint, _uint, _uint)
_sig("goofy_hip_compress_etc1", ctypes.c_int, _u8p, _u8p, _uint, _uint, _uint)
_sig("goofy_hip_compress_dxt5", ctypes.c_int, _u8p, _u8p, _uint, _uint, _uint)
_sig("goofy_hip_compress_etc2", ctypes.c_int, _u8p, _u8p, _uint, _uint, _uint)
_sig("goofy_format_block_bytes", _uint, ctypes.c_int)
_sig("goofy_hip_compress_device", ctypes.c_int, ctypes.c_int, _vp, _vp, _uint, _uint, _sz, _vp)
_sig("goofy_hip_compress_device_cfg", ctypes.c_int, ctypes.c_int, _vp, _vp, _uint, _uint, _sz, _vp,
     ctypes.POINTER(LaunchConfig))
_sig("goofy_hip_compress_device_timed", ctypes.c_int, ctypes.c_int, _vp, _vp, _uint, _uint, _sz, _vp,
     ctypes.POINTER(LaunchConfig), _vp, _vp)
_sig("goofy_batch_create", ctypes.c_int, ctypes.POINTER(TextureDesc), _sz, ctypes.POINTER(LaunchConfig),
     ctypes.POINTER(_vp))
_sig("goofy_batch_launch", ctypes.c_int, _vp, _vp)
_sig("goofy_batch_launch_per_texture", ctypes.c_int, _vp, _vp)
_sig("goofy_batch_autotune", ctypes.c_int, _vp, _vp, ctypes.c_int, ctypes.POINTER(LaunchConfig),
     ctypes.POINTER(ctypes.c_float))
_sig("goofy_batch_get_config", ctypes.c_int, _vp, ctypes.POINTER(LaunchConfig))
_sig("goofy_batch_set_config", ctypes.c_int, _vp, ctypes.POINTER(LaunchConfig))
_sig("goofy_batch_info", ctypes.c_int, _vp, ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_uint64),
     ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_uint64))
_sig("goofy_batch_destroy", ctypes.c_int, _vp)
_sig("goofy_synth_fill_host", ctypes.c_int, ctypes.c_int, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32, _u8p,
     _uint, _uint, _sz, ctypes.c_int)
_sig("goofy_synth_fill_device", ctypes.c_int, ctypes.c_int, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32, _vp,
     _uint, _uint, _sz, _vp)

# The reference-signature entry points, usable wherever a CompressFunc_t is.
compress_dxt1_fn = CompressFunc(("goofy_hip_compress_dxt1", _lib))
compress_etc1_fn = CompressFunc(("goofy_hip_compress_etc1", _lib))
compress_dxt5_fn = CompressFunc(("goofy_hip_compress_dxt5", _lib))
compress_etc2_fn = CompressFunc(("goofy_hip_compress_etc2", _lib))


def last_error() -> str:
    return _lib.goofy_hip_last_error().decode()


def version() -> str:
    return _lib.goofy_hip_version().decode()


def device_count() -> int:
    return int(_lib.goofy_hip_device_count())


def _check(status: int) -> int:
    if status != OK:
        raise GoofyError(status, last_error() or _lib.goofy_hip_status_string(status).decode())
    return status


def _u8(a) -> ctypes.POINTER(ctypes.c_uint8):
    if a is None:
        return ctypes.cast(None, _u8p)
    return a.ctypes.data_as(_u8p)


def _ptr(x) -> int | None:
    """Device address from an int or any object with data_ptr() (goofy_amd.hip.DeviceBuffer)."""
    if x is None:
        return None
    if isinstance(x, int):
        return x
    return int(x.data_ptr())


def _stream(s) -> int | None:
    """Stream handle from None (default stream), an int, or a goofy_amd.hip.Stream."""
    if s is None or isinstance(s, int):
        return s
    return int(s)


# ---- drop-in (host memory) ------------------------------------------------------

def compressDXT1(result: np.ndarray, input: np.ndarray, width: int, height: int, stride: int) -> int:
    """goofy::compressDXT1 (GoofyTC/goofy_tc.h:1497): returns 0 / -1 / -2 like the reference."""
    return int(_lib.goofy_hip_compress_dxt1(_u8(result), _u8(input), width, height, stride))


def compressETC1(result: np.ndarray, input: np.ndarray, width: int, height: int, stride: int) -> int:
    """goofy::compressETC1 (GoofyTC/goofy_tc.h:1528): returns 0 / -1 / -2 like the reference."""
    return int(_lib.goofy_hip_compress_etc1(_u8(result), _u8(input), width, height, stride))


def compressDXT5(result: np.ndarray, input: np.ndarray, width: int, height: int, stride: int) -> int:
    """DXT5 (BC3) with the drop-in's contract: 16 bytes per block, codes 0 / -1 / -2."""
    return int(_lib.goofy_hip_compress_dxt5(_u8(result), _u8(input), width, height, stride))


def compressETC2(result: np.ndarray, input: np.ndarray, width: int, height: int, stride: int) -> int:
    """ETC2 RGBA8 (EAC alpha) with the drop-in's contract: 16 bytes per block, codes 0 / -1 / -2."""
    return int(_lib.goofy_hip_compress_etc2(_u8(result), _u8(input), width, height, stride))


def compress(fmt: int, image: np.ndarray, width: int, height: int, stride: int | None = None) -> np.ndarray:
    """Encode a host RGBA8 image; returns the (w/4)*(h/4) blocks' bytes.  Raises on error."""
    stride = width * 4 if stride is None else stride
    n = (width // 4) * (height // 4) * block_bytes(fmt)
    out = np.empty(max(1, n), dtype=np.uint8)
    _check(_lib.goofy_hip_compress(fmt, _u8(out), _u8(image), width, height, stride))
    return out[:n]


_sig("goofy_hip_compress_dual", ctypes.c_int, _u8p, _u8p, _u8p, _uint, _uint, _uint)


_sig("goofy_hip_compress_dual_alpha", ctypes.c_int, _u8p, _u8p, _u8p, _uint, _uint, _uint)


def compress_dual_alpha(image: np.ndarray, width: int, height: int, stride: int | None = None) -> tuple:
    """DXT5 and ETC2 RGBA8 of a host RGBA8 image from one transfer and one pass: (dxt5, etc2) block bytes."""
    stride = width * 4 if stride is None else stride
    n = (width // 4) * (height // 4) * 16
    a = np.empty(max(1, n), dtype=np.uint8)
    b = np.empty(max(1, n), dtype=np.uint8)
    _check(_lib.goofy_hip_compress_dual_alpha(_u8(a), _u8(b), _u8(image), width, height, stride))
    return a[:n], b[:n]


def compress_dual(image: np.ndarray, width: int, height: int, stride: int | None = None, out=None) -> tuple:
    """Both formats of a host RGBA8 image from one transfer and one pass: (dxt1, etc1s) block bytes.
    `out`: an optional pair of uint8 arrays of at least (w/4)*(h/4)*8 bytes each to write into."""
    stride = width * 4 if stride is None else stride
    n = (width // 4) * (height // 4) * 8
    if out is None:
        a = np.empty(max(1, n), dtype=np.uint8)
        b = np.empty(max(1, n), dtype=np.uint8)
    else:
        a, b = out
        if a.dtype != np.uint8 or b.dtype != np.uint8 or a.size < n or b.size < n or not (
                a.flags.c_contiguous and b.flags.c_contiguous):
            raise ValueError(f"out needs two contiguous uint8 arrays of >= {n} bytes")
    _check(_lib.goofy_hip_compress_dual(_u8(a), _u8(b), _u8(image), width, height, stride))
    return a[:n], b[:n]


# ---- device memory ----------------------------------------------------------------

def compress_device(fmt: int, dst, src, width: int, height: int, pitch: int, stream=None,
                    config: LaunchConfig | dict | None = None) -> None:
    """Enqueue the encode of a device-resident image (dst/src: device addresses or DeviceBuffers)."""
    if config is None:
        _check(_lib.goofy_hip_compress_device(fmt, _ptr(dst), _ptr(src), width, height, pitch, _stream(stream)))
    else:
        cfg = config if isinstance(config, LaunchConfig) else LaunchConfig(**config)
        _check(_lib.goofy_hip_compress_device_cfg(fmt, _ptr(dst), _ptr(src), width, height, pitch,
                                                  _stream(stream), ctypes.byref(cfg)))


_sig("goofy_hip_compress_device_dual", ctypes.c_int, _vp, _vp, _vp, _uint, _uint, _sz, _vp, _vp, _vp)


def compress_device_dual(dst_dxt1, dst_etc1, src, width: int, height: int, pitch: int, stream=None,
                         start_event=None, stop_event=None) -> None:
    """Both formats of a device-resident image in one pass (one read of the image)."""
    _check(_lib.goofy_hip_compress_device_dual(_ptr(dst_dxt1), _ptr(dst_etc1), _ptr(src), width, height, pitch,
                                               _stream(stream), start_event.handle if start_event else None,
                                               stop_event.handle if stop_event else None))


_sig("goofy_hip_compress_device_dual_alpha", ctypes.c_int, _vp, _vp, _vp, _uint, _uint, _sz, _vp, _vp, _vp)


def compress_device_dual_alpha(dst_dxt5, dst_etc2, src, width: int, height: int, pitch: int, stream=None,
                               start_event=None, stop_event=None) -> None:
    """DXT5 and ETC2 RGBA8 of a device-resident image in one pass (one read of the image)."""
    _check(_lib.goofy_hip_compress_device_dual_alpha(_ptr(dst_dxt5), _ptr(dst_etc2), _ptr(src), width, height, pitch,
                                                     _stream(stream), start_event.handle if start_event else None,
                                                     stop_event.handle if stop_event else None))


_sig("goofy_hip_tune_device", ctypes.c_int, ctypes.c_int, _vp, _vp, _uint, _uint, _sz, _vp, ctypes.c_int,
     ctypes.POINTER(LaunchConfig), ctypes.POINTER(ctypes.c_float))


def tune_device(fmt: int, dst, src, width: int, height: int, pitch: int, stream=None, reps: int = 3) -> tuple:
    """Fastest launch shape for this image shape: (config dict, median kernel ms)."""
    cfg = LaunchConfig()
    ms = ctypes.c_float()
    _check(_lib.goofy_hip_tune_device(fmt, _ptr(dst), _ptr(src), width, height, pitch, _stream(stream), reps,
                                      ctypes.byref(cfg), ctypes.byref(ms)))
    return cfg.as_dict(), float(ms.value)


def compress_device_timed(fmt: int, dst, src, width: int, height: int, pitch: int, start_event, stop_event,
                          stream=None, config: LaunchConfig | dict | None = None) -> None:
    """compress_device with hipEvents recorded by the kernel's own dispatch (goofy_amd.hip.Event)."""
    cfg = None
    if config is not None:
        cfg = config if isinstance(config, LaunchConfig) else LaunchConfig(**config)
    _check(_lib.goofy_hip_compress_device_timed(fmt, _ptr(dst), _ptr(src), width, height, pitch, _stream(stream),
                                                ctypes.byref(cfg) if cfg else None,
                                                start_event.handle if start_event else None,
                                                stop_event.handle if stop_event else None))


class Quality(ctypes.Structure):
    _fields_ = ([("pixels", ctypes.c_uint64), ("sse_byte", ctypes.c_uint64 * 4), ("sse_luma", ctypes.c_double)] +
                [(n, ctypes.c_double) for n in ("mse_r", "mse_g", "mse_b", "mse_a", "mse_max", "mse_rgb", "mse_y",
                                                "psnr_r", "psnr_g", "psnr_b", "psnr_a", "psnr_min", "psnr_rgb",
                                                "psnr_y")])

    def as_dict(self) -> dict:
        d = {n: getattr(self, n) for n, _ in self._fields_}
        d["sse_byte"] = list(self.sse_byte)
        return d


_sig("goofy_hip_decode_device", ctypes.c_int, ctypes.c_int, _vp, _vp, _uint, _uint, _sz, _vp)
_sig("goofy_hip_quality_device", ctypes.c_int, ctypes.c_int, _vp, _vp, _uint, _uint, _sz, _vp,
     ctypes.POINTER(Quality))


def decode_device(fmt: int, dst_rgba, blocks, width: int, height: int, pitch: int | None = None,
                  stream=None) -> None:
    """Decode device-resident blocks into RGBA8 rows of `pitch` bytes (HIP kernel)."""
    pitch = width * 4 if pitch is None else pitch
    _check(_lib.goofy_hip_decode_device(fmt, _ptr(dst_rgba), _ptr(blocks), width, height, pitch, _stream(stream)))


def quality_device(fmt: int, blocks, source, width: int, height: int, pitch: int | None = None,
                   stream=None) -> dict:
    """Round-trip MSE/PSNR of encoded blocks against the source image (definitions of
    the reference harness's getMsePsnr, Src/main.cpp:403-469)."""
    pitch = width * 4 if pitch is None else pitch
    q = Quality()
    _check(_lib.goofy_hip_quality_device(fmt, _ptr(blocks), _ptr(source), width, height, pitch, _stream(stream),
                                         ctypes.byref(q)))
    return q.as_dict()


_sig("goofy_hip_quality_device_noalloc", ctypes.c_int, ctypes.c_int, _vp, _vp, _uint, _uint, _sz, _vp,
     ctypes.POINTER(Quality))


def quality_device_noalloc(fmt: int, blocks, source, width: int, height: int, pitch: int | None = None,
                           stream=None) -> dict:
    """quality_device without allocating: the library's fixed per-GPU device
    array is the work area.  Calls that overlap on one GPU race."""
    pitch = width * 4 if pitch is None else pitch
    q = Quality()
    _check(_lib.goofy_hip_quality_device_noalloc(fmt, _ptr(blocks), _ptr(source), width, height, pitch,
                                                 _stream(stream), ctypes.byref(q)))
    return q.as_dict()


_sig("goofy_hip_quality_images_device", ctypes.c_int, _vp, _sz, _vp, _sz, _uint, _uint, _vp,
     ctypes.POINTER(Quality))
_sig("goofy_hip_rgb565_baseline_device", ctypes.c_int, _vp, _sz, _vp, _sz, _uint, _uint, _vp)
_sig("goofy_hip_rgb565_baseline_device_timed", ctypes.c_int, _vp, _sz, _vp, _sz, _uint, _uint, _vp, _vp, _vp)


def quality_images_device(a, b, width: int, height: int, pitch_a: int | None = None, pitch_b: int | None = None,
                          stream=None) -> dict:
    """getMsePsnr figures (Src/main.cpp:403-469) of two device RGBA8 images."""
    pitch_a = width * 4 if pitch_a is None else pitch_a
    pitch_b = width * 4 if pitch_b is None else pitch_b
    q = Quality()
    _check(_lib.goofy_hip_quality_images_device(_ptr(a), pitch_a, _ptr(b), pitch_b, width, height, _stream(stream),
                                                ctypes.byref(q)))
    return q.as_dict()


def rgb565_baseline_device(dst, src, width: int, height: int, dst_pitch: int | None = None,
                           src_pitch: int | None = None, stream=None, start_event=None, stop_event=None) -> None:
    """The harness's quality floor, "Baseline (Downsample x2)" in RGB565
    (generateDownsamlpedRgb565Test, Src/main.cpp:501-555), on the GPU."""
    dst_pitch = width * 4 if dst_pitch is None else dst_pitch
    src_pitch = width * 4 if src_pitch is None else src_pitch
    if start_event is None and stop_event is None:
        _check(_lib.goofy_hip_rgb565_baseline_device(_ptr(dst), dst_pitch, _ptr(src), src_pitch, width, height,
                                                     _stream(stream)))
    else:
        _check(_lib.goofy_hip_rgb565_baseline_device_timed(
            _ptr(dst), dst_pitch, _ptr(src), src_pitch, width, height, _stream(stream),
            start_event.handle if start_event else None, stop_event.handle if stop_event else None))


_sig("goofy_hip_traffic_probe", ctypes.c_int, _vp, _vp, _uint, _uint, _sz, ctypes.c_int, _vp, _vp, _vp)


def traffic_probe(dst, src, width: int, height: int, pitch: int, start_event=None, stop_event=None,
                  stream=None, wg_waves: int = 0) -> None:
    """The encode kernel's memory traffic without its arithmetic (roofline calibration)."""
    _check(_lib.goofy_hip_traffic_probe(_ptr(dst), _ptr(src), width, height, pitch, wg_waves, _stream(stream),
                                        start_event.handle if start_event else None,
                                        stop_event.handle if stop_event else None))


class Batch:
    """Many device-resident textures encoded by one persistent-kernel launch (hipGraph replay)."""

    def __init__(self, textures: Sequence[dict], config: LaunchConfig | dict | None = None):
        arr = (TextureDesc * max(1, len(textures)))()
        for i, t in enumerate(textures):
            arr[i] = TextureDesc(_ptr(t["src"]), _ptr(t["dst"]), t["width"], t["height"],
                                 t.get("pitch", t["width"] * 4), t["format"])
        h = _vp()
        cfg = None
        if config is not None:
            cfg = config if isinstance(config, LaunchConfig) else LaunchConfig(**config)
        _check(_lib.goofy_batch_create(arr, len(textures), ctypes.byref(cfg) if cfg else None, ctypes.byref(h)))
        self._h = h
        self._keep = arr

    def launch(self, stream=None) -> None:
        _check(_lib.goofy_batch_launch(self._h, _stream(stream)))

    def launch_per_texture(self, stream=None) -> None:
        _check(_lib.goofy_batch_launch_per_texture(self._h, _stream(stream)))

    def autotune(self, stream=None, reps: int = 3) -> tuple[dict, float]:
        cfg = LaunchConfig()
        ms = ctypes.c_float()
        _check(_lib.goofy_batch_autotune(self._h, _stream(stream), reps, ctypes.byref(cfg), ctypes.byref(ms)))
        return cfg.as_dict(), float(ms.value)

    @property
    def config(self) -> dict:
        cfg = LaunchConfig()
        _check(_lib.goofy_batch_get_config(self._h, ctypes.byref(cfg)))
        return cfg.as_dict()

    @config.setter
    def config(self, value: dict) -> None:
        cfg = LaunchConfig(**value)
        _check(_lib.goofy_batch_set_config(self._h, ctypes.byref(cfg)))

    def info(self) -> dict:
        v = [ctypes.c_uint64() for _ in range(4)]
        _check(_lib.goofy_batch_info(self._h, *[ctypes.byref(x) for x in v]))
        return dict(zip(("textures", "blocks", "tiles", "pixels"), (int(x.value) for x in v)))

    def close(self) -> None:
        if getattr(self, "_h", None):
            _lib.goofy_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---- synthetic inputs -----------------------------------------------------------------

WAVES, STRESS = 0, 1


def synth_host(kind: int, seed: int, texture_id: int, noise: int, width: int, height: int,
               pitch: int | None = None, threads: int = 0, fill: int = 0) -> np.ndarray:
    """Seeded synthetic RGBA8 texture in host memory; returns a flat uint8 array of pitch*height
    bytes (256-byte aligned).  Padding bytes beyond 4*width hold `fill`."""
    pitch = width * 4 if pitch is None else pitch
    raw = np.empty(pitch * height + 256, dtype=np.uint8)
    off = (-raw.ctypes.data) % 256
    img = raw[off:off + pitch * height]
    if pitch != width * 4:
        img.fill(fill)
    _check(_lib.goofy_synth_fill_host(kind, seed, texture_id, noise, _u8(img), width, height, pitch, threads))
    return img


def synth_device(kind: int, seed: int, texture_id: int, noise: int, dst, width: int, height: int,
                 pitch: int | None = None, stream=None) -> None:
    pitch = width * 4 if pitch is None else pitch
    _check(_lib.goofy_synth_fill_device(kind, seed, texture_id, noise, _ptr(dst), width, height, pitch,
                                        _stream(stream)))


__all__ = [
    "DXT1", "ETC1", "DXT5", "ETC2", "FORMATS", "block_bytes", "compressDXT5", "compressETC2",
    "compress_dxt5_fn", "compress_etc2_fn", "GoofyError", "LaunchConfig", "TextureDesc", "CompressFunc", "Batch",
    "compressDXT1", "compressETC1", "compress", "compress_device", "compress_dxt1_fn", "compress_etc1_fn",
    "synth_host", "synth_device", "last_error", "version", "device_count", "WAVES", "STRESS",
    "decode_device", "quality_device", "quality_device_noalloc", "compress_device_timed", "quality_images_device", "rgb565_baseline_device",
    "tune_device", "compress_device_dual",
]
