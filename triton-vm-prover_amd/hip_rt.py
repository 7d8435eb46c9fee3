"""Minimal HIP runtime plumbing for Python callers of libtvh.

Device memory, streams and events come from the same libamdhip64 that libtvh.so
links (/opt/rocm), loaded through ctypes, so a process holds exactly one HIP
runtime.  (Mixing in a second runtime, e.g. the one bundled with a PyTorch
wheel, gives two independent HIP states in one process.)
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HIP = None


def hip():
    global _HIP
    if _HIP is None:
        path = os.environ.get("TVH_HIP_RUNTIME", "/opt/rocm/lib/libamdhip64.so")
        lib = C.CDLL(path, mode=C.RTLD_GLOBAL)
        v, sz, i = C.c_void_p, C.c_size_t, C.c_int
        sigs = {
            "hipMalloc": [C.POINTER(v), sz], "hipFree": [v],
            "hipMemcpy": [v, v, sz, i], "hipMemcpyAsync": [v, v, sz, i, v],
            "hipMemcpy2D": [v, sz, v, sz, sz, sz, i],
            "hipMemset": [v, i, sz], "hipMemsetAsync": [v, i, sz, v],
            "hipDeviceSynchronize": [], "hipGetDeviceCount": [C.POINTER(i)],
            "hipSetDevice": [i], "hipGetLastError": [],
            "hipStreamCreateWithFlags": [C.POINTER(v), C.c_uint],
            "hipStreamDestroy": [v], "hipStreamSynchronize": [v],
            "hipEventCreate": [C.POINTER(v)], "hipEventDestroy": [v],
            "hipEventRecord": [v, v], "hipEventSynchronize": [v],
            "hipEventElapsedTime": [C.POINTER(C.c_float), v, v],
            "hipMemGetInfo": [C.POINTER(sz), C.POINTER(sz)],
            "hipGetErrorString": [i],
            "hipHostMalloc": [C.POINTER(v), sz, C.c_uint], "hipHostFree": [v],
            "hipStreamWaitEvent": [v, v, C.c_uint],
        }
        for name, args in sigs.items():
            f = getattr(lib, name)
            f.argtypes = args
            f.restype = C.c_char_p if name == "hipGetErrorString" else C.c_int
        _HIP = _Runtime(lib, sigs)
    return _HIP


class _Runtime:
    """The HIP runtime's declared entry points only.  ctypes passes the
    arguments of an undeclared function as 32-bit ints, which truncates
    pointers and sizes silently, so any other name raises instead."""

    def __init__(self, lib, sigs):
        self._lib = lib  # keeps the library loaded
        self._fns = {name: getattr(lib, name) for name in sigs}

    def __getattr__(self, name):
        try:
            return self.__dict__["_fns"][name]
        except KeyError:
            raise AttributeError(f"HIP runtime function {name} has no declared signature") from None


H2D, D2H, D2D = 1, 2, 3


class HipError(RuntimeError):
    pass


def check(rc: int, what: str) -> None:
    if rc:
        raise HipError(f"{what}: {hip().hipGetErrorString(rc).decode()} ({rc})")


def device_count() -> int:
    n = C.c_int(0)
    rc = hip().hipGetDeviceCount(C.byref(n))
    return n.value if rc == 0 else 0


def synchronize() -> None:
    check(hip().hipDeviceSynchronize(), "hipDeviceSynchronize")


def mem_info() -> tuple[int, int]:
    free, total = C.c_size_t(0), C.c_size_t(0)
    check(hip().hipMemGetInfo(C.byref(free), C.byref(total)), "hipMemGetInfo")
    return free.value, total.value


class DeviceBuffer:
    """u64 elements in device memory (owning, or a view at an element offset)."""

    def __init__(self, count: int, _base=None, _offset: int = 0):
        self.count = int(count)
        if _base is None:
            p = C.c_void_p()
            check(hip().hipMalloc(C.byref(p), max(8, 8 * self.count)), f"hipMalloc {8 * self.count} B")
            self._raw = p.value
            self._owner = True
            self.ptr = p.value
        else:
            self._raw = None
            self._owner = False
            self._base = _base  # keep the owner alive
            self.ptr = _base.ptr + 8 * _offset

    def data_ptr(self) -> int:
        return self.ptr

    def view(self, offset: int, count: int) -> "DeviceBuffer":
        if offset < 0 or offset + count > self.count:
            raise IndexError("view out of range")
        return DeviceBuffer(count, _base=self, _offset=offset)

    def upload(self, a: np.ndarray, stream=None) -> "DeviceBuffer":
        a = np.ascontiguousarray(a, dtype=np.uint64).reshape(-1)
        if a.size > self.count:
            raise ValueError("upload larger than buffer")
        if a.size:
            if stream is None:
                check(hip().hipMemcpy(self.ptr, a.ctypes.data, 8 * a.size, H2D), "H2D")
            else:
                check(hip().hipMemcpyAsync(self.ptr, a.ctypes.data, 8 * a.size, H2D, stream.ptr),
                      "H2D async")
                stream.synchronize()
        return self

    def download(self, count: int | None = None, offset: int = 0) -> np.ndarray:
        count = self.count - offset if count is None else count
        out = np.empty(count, np.uint64)
        if count:
            check(hip().hipMemcpy(out.ctypes.data, self.ptr + 8 * offset, 8 * count, D2H), "D2H")
        return out

    def gather_strided(self, offset: int, stride: int, count: int) -> np.ndarray:
        """count elements at offset, offset + stride, ... (a table row of a column-major table)."""
        out = np.empty(count, np.uint64)
        check(hip().hipMemcpy2D(out.ctypes.data, 8, self.ptr + 8 * offset, 8 * stride, 8, count, D2H),
              "hipMemcpy2D")
        return out

    def zero(self) -> "DeviceBuffer":
        check(hip().hipMemset(self.ptr, 0, 8 * self.count), "hipMemset")
        return self

    def free(self) -> None:
        if self._owner and self._raw:
            hip().hipFree(self._raw)
            self._raw = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def to_device(a: np.ndarray) -> DeviceBuffer:
    a = np.ascontiguousarray(a, dtype=np.uint64).reshape(-1)
    return DeviceBuffer(max(1, a.size)).upload(a)


class Stream:
    def __init__(self, nonblocking: bool = True):
        s = C.c_void_p()
        check(hip().hipStreamCreateWithFlags(C.byref(s), 1 if nonblocking else 0), "stream create")
        self.ptr = s.value

    def synchronize(self) -> None:
        check(hip().hipStreamSynchronize(self.ptr), "hipStreamSynchronize")

    def destroy(self) -> None:
        if self.ptr:
            hip().hipStreamDestroy(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


class Event:
    def __init__(self):
        e = C.c_void_p()
        check(hip().hipEventCreate(C.byref(e)), "event create")
        self.ptr = e.value

    def record(self, stream: Stream | None = None) -> None:
        check(hip().hipEventRecord(self.ptr, stream.ptr if stream else None), "event record")

    def synchronize(self) -> None:
        check(hip().hipEventSynchronize(self.ptr), "event sync")

    def elapsed_ms(self, end: "Event") -> float:
        ms = C.c_float(0)
        check(hip().hipEventElapsedTime(C.byref(ms), self.ptr, end.ptr), "elapsed")
        return ms.value

    def __del__(self):
        try:
            if self.ptr:
                hip().hipEventDestroy(self.ptr)
        except Exception:
            pass
