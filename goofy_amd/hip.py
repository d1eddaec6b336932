"""Minimal ctypes access to the HIP runtime that libgoofy_hip.so links
(device buffers, streams, events) -- plumbing for tests, the benchmark and
Python callers that have no other device-memory library.

It binds the very same libamdhip64 instance as the encoder library (loaded
by soname after the library), so pointers, streams and events are shared
with the C-ABI.  Do not mix with a second HIP runtime in the same process
(e.g. the copy bundled inside a PyTorch wheel).
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib  # noqa: F401  (loads libgoofy_hip.so and, through it, libamdhip64.so.7)

_hip = ctypes.CDLL("libamdhip64.so.7")
_vp = ctypes.c_void_p
_sz = ctypes.c_size_t

H2D, D2H, D2D = 1, 2, 3


def _sig(name, *argtypes, restype=ctypes.c_int):
    fn = getattr(_hip, name)
    fn.restype = restype
    fn.argtypes = list(argtypes)
    return fn


_sig("hipGetErrorString", ctypes.c_int, restype=ctypes.c_char_p)
_sig("hipGetDeviceCount", ctypes.POINTER(ctypes.c_int))
_sig("hipSetDevice", ctypes.c_int)
_sig("hipMalloc", ctypes.POINTER(_vp), _sz)
_sig("hipFree", _vp)
_sig("hipHostMalloc", ctypes.POINTER(_vp), _sz, ctypes.c_uint)
_sig("hipHostFree", _vp)
_sig("hipMemcpy", _vp, _vp, _sz, ctypes.c_int)
_sig("hipMemcpyAsync", _vp, _vp, _sz, ctypes.c_int, _vp)
_sig("hipMemset", _vp, ctypes.c_int, _sz)
_sig("hipDeviceSynchronize")
_sig("hipStreamCreateWithFlags", ctypes.POINTER(_vp), ctypes.c_uint)
_sig("hipStreamDestroy", _vp)
_sig("hipStreamSynchronize", _vp)
_sig("hipEventCreate", ctypes.POINTER(_vp))
_sig("hipEventDestroy", _vp)
_sig("hipEventRecord", _vp, _vp)
_sig("hipEventSynchronize", _vp)
_sig("hipEventElapsedTime", ctypes.POINTER(ctypes.c_float), _vp, _vp)


class HipError(RuntimeError):
    pass


def check(rc: int, what: str = "") -> None:
    if rc != 0:
        raise HipError(f"{what}: HIP error {rc} ({_hip.hipGetErrorString(rc).decode()})")


def device_count() -> int:
    n = ctypes.c_int(0)
    if _hip.hipGetDeviceCount(ctypes.byref(n)) != 0:
        return 0
    return n.value


def synchronize() -> None:
    check(_hip.hipDeviceSynchronize(), "hipDeviceSynchronize")


def _u8ptr(a: np.ndarray) -> int:
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data


class DeviceBuffer:
    """A device allocation (or a view into one) of `nbytes` bytes, aligned to `align`."""

    def __init__(self, nbytes: int, align: int = 256, fill: int | None = None, _base=None, _ptr: int = 0):
        self.nbytes = int(nbytes)
        self._owner = _base
        if _base is None:
            raw = _vp()
            check(_hip.hipMalloc(ctypes.byref(raw), max(1, self.nbytes + align)), f"hipMalloc({nbytes})")
            self._raw = raw.value
            self.ptr = self._raw + ((-self._raw) % align)
        else:
            self._raw = None
            self.ptr = _ptr
        if fill is not None:
            self.fill(fill)

    def data_ptr(self) -> int:
        return self.ptr

    def __int__(self) -> int:
        return self.ptr

    def view(self, offset: int, nbytes: int) -> "DeviceBuffer":
        assert 0 <= offset and offset + nbytes <= self.nbytes
        return DeviceBuffer(nbytes, _base=self, _ptr=self.ptr + offset)

    def fill(self, byte: int) -> "DeviceBuffer":
        if self.nbytes:
            check(_hip.hipMemset(self.ptr, int(byte), self.nbytes), "hipMemset")
        return self

    def upload(self, a: np.ndarray, offset: int = 0) -> "DeviceBuffer":
        a = np.ascontiguousarray(a)
        assert offset + a.nbytes <= self.nbytes
        if a.nbytes:
            check(_hip.hipMemcpy(self.ptr + offset, _u8ptr(a), a.nbytes, H2D), "hipMemcpy H2D")
        return self

    def download(self, nbytes: int | None = None, offset: int = 0) -> np.ndarray:
        n = self.nbytes - offset if nbytes is None else nbytes
        out = np.empty(n, dtype=np.uint8)
        if n:
            check(_hip.hipMemcpy(_u8ptr(out), self.ptr + offset, n, D2H), "hipMemcpy D2H")
        return out

    def copy_from(self, other: "DeviceBuffer") -> None:
        n = min(self.nbytes, other.nbytes)
        check(_hip.hipMemcpy(self.ptr, other.ptr, n, D2D), "hipMemcpy D2D")

    @classmethod
    def from_array(cls, a: np.ndarray, align: int = 256) -> "DeviceBuffer":
        b = cls(np.asarray(a).nbytes, align)
        return b.upload(a)

    def free(self) -> None:
        if self._raw is not None:
            _hip.hipFree(self._raw)
            self._raw = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Stream:
    def __init__(self):
        s = _vp()
        check(_hip.hipStreamCreateWithFlags(ctypes.byref(s), 1), "hipStreamCreate")  # non-blocking
        self.handle = s.value

    def synchronize(self) -> None:
        check(_hip.hipStreamSynchronize(self.handle), "hipStreamSynchronize")

    def __int__(self) -> int:
        return self.handle

    def destroy(self) -> None:
        if self.handle:
            _hip.hipStreamDestroy(self.handle)
            self.handle = None


class Event:
    def __init__(self):
        e = _vp()
        check(_hip.hipEventCreate(ctypes.byref(e)), "hipEventCreate")
        self.handle = e.value

    def record(self, stream=None) -> "Event":
        check(_hip.hipEventRecord(self.handle, int(stream) if stream is not None else None), "hipEventRecord")
        return self

    def synchronize(self) -> None:
        check(_hip.hipEventSynchronize(self.handle), "hipEventSynchronize")

    def elapsed_ms(self, end: "Event") -> float:
        ms = ctypes.c_float()
        check(_hip.hipEventElapsedTime(ctypes.byref(ms), self.handle, end.handle), "hipEventElapsedTime")
        return float(ms.value)

    def __del__(self):
        try:
            if self.handle:
                _hip.hipEventDestroy(self.handle)
                self.handle = None
        except Exception:
            pass
