"""HIP streams and events over ctypes, for the serving loop and the bench.

The HIP runtime is the one libnewt_gpu.so links against (libamdhip64.so.7);
loading the library first makes the dynamic loader hand back that same
runtime here, so the streams and events below are valid handles for
newt_step and the newt_data_* calls.  Every stream is created at the default
priority.
"""
from __future__ import annotations

import ctypes as C

from . import abi

_hip = None


def runtime():
    global _hip
    if _hip is None:
        abi.load_library()
        h = C.CDLL("libamdhip64.so.7")
        P, I = C.c_void_p, C.c_int
        sig = {
            "hipStreamCreateWithFlags": (I, [C.POINTER(P), C.c_uint]),
            "hipStreamDestroy": (I, [P]),
            "hipStreamSynchronize": (I, [P]),
            "hipEventCreateWithFlags": (I, [C.POINTER(P), C.c_uint]),
            "hipEventDestroy": (I, [P]),
            "hipEventRecord": (I, [P, P]),
            "hipEventQuery": (I, [P]),
            "hipEventSynchronize": (I, [P]),
            "hipEventElapsedTime": (I, [C.POINTER(C.c_float), P, P]),
            "hipDeviceSynchronize": (I, []),
            "hipGetDeviceCount": (I, [C.POINTER(I)]),
            "hipGetErrorString": (C.c_char_p, [I]),
            "hipMalloc": (I, [C.POINTER(P), C.c_size_t]),
            "hipFree": (I, [P]),
            "hipMemcpy": (I, [P, P, C.c_size_t, I]),
        }
        for name, (res, args) in sig.items():
            f = getattr(h, name)
            f.restype = res
            f.argtypes = args
        _hip = h
    return _hip


def _check(err, what):
    if err != 0:
        raise RuntimeError(f"{what}: {runtime().hipGetErrorString(err).decode()}")


HIP_ERROR_NOT_READY = 600
STREAM_NON_BLOCKING = 1
EVENT_DISABLE_TIMING = 2


def device_count() -> int:
    n = C.c_int(0)
    if runtime().hipGetDeviceCount(C.byref(n)) != 0:
        return 0
    return n.value


def synchronize():
    _check(runtime().hipDeviceSynchronize(), "hipDeviceSynchronize")


class Stream:
    """A non-blocking HIP stream at the default priority."""

    def __init__(self):
        h = C.c_void_p()
        _check(runtime().hipStreamCreateWithFlags(C.byref(h), STREAM_NON_BLOCKING), "hipStreamCreateWithFlags")
        self.handle = h.value

    def synchronize(self):
        _check(runtime().hipStreamSynchronize(self.handle), "hipStreamSynchronize")

    def close(self):
        if self.handle:
            runtime().hipStreamDestroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Event:
    def __init__(self, timing: bool = True):
        h = C.c_void_p()
        _check(runtime().hipEventCreateWithFlags(C.byref(h), 0 if timing else EVENT_DISABLE_TIMING),
               "hipEventCreateWithFlags")
        self.handle = h.value

    def record(self, stream: Stream | None = None):
        _check(runtime().hipEventRecord(self.handle, stream.handle if stream else None), "hipEventRecord")

    def query(self) -> bool:
        err = runtime().hipEventQuery(self.handle)
        if err == HIP_ERROR_NOT_READY:
            return False
        _check(err, "hipEventQuery")
        return True

    def synchronize(self):
        _check(runtime().hipEventSynchronize(self.handle), "hipEventSynchronize")

    def elapsed_ms(self, end: "Event") -> float:
        ms = C.c_float()
        _check(runtime().hipEventElapsedTime(C.byref(ms), self.handle, end.handle), "hipEventElapsedTime")
        return float(ms.value)

    def close(self):
        if self.handle:
            runtime().hipEventDestroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


HOST_TO_DEVICE, DEVICE_TO_HOST = 1, 2


class DeviceArray:
    """A device copy of a numpy array (hipMalloc'd; freed by close())."""

    def __init__(self, host):
        import numpy as np
        host = np.ascontiguousarray(host)
        self.shape, self.dtype = host.shape, host.dtype
        self.nbytes = max(host.nbytes, 8)
        h = C.c_void_p()
        _check(runtime().hipMalloc(C.byref(h), self.nbytes), "hipMalloc")
        self.ptr = h.value
        self.upload(host)

    def upload(self, host):
        import numpy as np
        host = np.ascontiguousarray(host, dtype=self.dtype)
        if host.nbytes:
            _check(runtime().hipMemcpy(self.ptr, host.ctypes.data, host.nbytes, HOST_TO_DEVICE), "hipMemcpy")

    def download(self):
        import numpy as np
        out = np.empty(self.shape, dtype=self.dtype)
        if out.nbytes:
            _check(runtime().hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, DEVICE_TO_HOST), "hipMemcpy")
        return out

    def close(self):
        if self.ptr:
            runtime().hipFree(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
