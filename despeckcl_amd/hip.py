"""Minimal HIP runtime access through ctypes: device buffers, streams and events.

Used by the device entry points' callers that do not bring their own device memory
(benchmark, smoke run, serving loop).  Everything here is plumbing around the HIP runtime;
all filtering happens in libdespeckcl_amd.so.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_hip = None


def hip():
    global _hip
    if _hip is None:
        for name in ("libamdhip64.so", "libamdhip64.so.7", "/opt/rocm/lib/libamdhip64.so"):
            try:
                _hip = C.CDLL(name)
                break
            except OSError:
                continue
        if _hip is None:
            raise OSError("libamdhip64.so not found")
        _hip.hipGetErrorString.restype = C.c_char_p
        _hip.hipEventElapsedTime.argtypes = [C.POINTER(C.c_float), C.c_void_p, C.c_void_p]
    return _hip


def _check(err, what):
    if err != 0:
        raise RuntimeError(f"{what} failed: {hip().hipGetErrorString(err).decode()}")


def device_count():
    n = C.c_int(0)
    err = hip().hipGetDeviceCount(C.byref(n))
    return n.value if err == 0 else 0


def synchronize():
    _check(hip().hipDeviceSynchronize(), "hipDeviceSynchronize")


class DeviceArray:
    """A float32 device buffer of a given shape."""

    def __init__(self, shape, dtype=np.float32):
        self.shape = tuple(shape)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = C.c_void_p()
        _check(hip().hipMalloc(C.byref(p), C.c_size_t(max(self.nbytes, 4))), "hipMalloc")
        self.ptr = p.value

    @classmethod
    def from_host(cls, a, stream=None):
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype)
        d.upload(a, stream)
        return d

    def upload(self, a, stream=None):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.nbytes == self.nbytes
        if stream is None:
            _check(hip().hipMemcpy(C.c_void_p(self.ptr), a.ctypes.data_as(C.c_void_p),
                                   C.c_size_t(self.nbytes), 1), "hipMemcpy H2D")
        else:
            _check(hip().hipMemcpyAsync(C.c_void_p(self.ptr), a.ctypes.data_as(C.c_void_p),
                                        C.c_size_t(self.nbytes), 1, C.c_void_p(stream)), "hipMemcpyAsync")

    def download(self, out=None):
        out = np.empty(self.shape, self.dtype) if out is None else out
        _check(hip().hipMemcpy(out.ctypes.data_as(C.c_void_p), C.c_void_p(self.ptr),
                               C.c_size_t(self.nbytes), 2), "hipMemcpy D2H")
        return out

    def zero(self):
        _check(hip().hipMemset(C.c_void_p(self.ptr), 0, C.c_size_t(self.nbytes)), "hipMemset")

    def free(self):
        if self.ptr:
            hip().hipFree(C.c_void_p(self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def copy2d(dst, dst_pitch_bytes, src, src_pitch_bytes, width_bytes, height):
    """Device-to-device copy of a height x width_bytes region between pitched buffers."""
    _check(hip().hipMemcpy2D(C.c_void_p(dst), C.c_size_t(dst_pitch_bytes), C.c_void_p(src),
                             C.c_size_t(src_pitch_bytes), C.c_size_t(width_bytes), C.c_size_t(height), 3),
           "hipMemcpy2D")


class Stream:
    def __init__(self):
        s = C.c_void_p()
        _check(hip().hipStreamCreate(C.byref(s)), "hipStreamCreate")
        self.handle = s.value

    def synchronize(self):
        _check(hip().hipStreamSynchronize(C.c_void_p(self.handle)), "hipStreamSynchronize")

    def destroy(self):
        if self.handle:
            hip().hipStreamDestroy(C.c_void_p(self.handle))
            self.handle = None


class Event:
    def __init__(self):
        e = C.c_void_p()
        _check(hip().hipEventCreate(C.byref(e)), "hipEventCreate")
        self.handle = e.value

    def record(self, stream=None):
        _check(hip().hipEventRecord(C.c_void_p(self.handle), C.c_void_p(stream or 0)), "hipEventRecord")

    def synchronize(self):
        _check(hip().hipEventSynchronize(C.c_void_p(self.handle)), "hipEventSynchronize")

    def elapsed_ms(self, end):
        ms = C.c_float()
        _check(hip().hipEventElapsedTime(C.byref(ms), C.c_void_p(self.handle), C.c_void_p(end.handle)),
               "hipEventElapsedTime")
        return ms.value


def device_name():
    class Props(C.Structure):
        _fields_ = [("name", C.c_char * 256), ("pad", C.c_char * 4096)]
    p = Props()
    if hip().hipGetDevicePropertiesR0600(C.byref(p), 0) != 0:
        return "unknown"
    return p.name.decode(errors="replace")


def env_summary():
    return {"device": device_name() if device_count() else None,
            "hip_visible": os.environ.get("HIP_VISIBLE_DEVICES")}
