"""Streams and events from the HIP runtime (ctypes on libamdhip64), for
callers that time or overlap work without a framework."""
import ctypes
import os

_hip = None


def hip():
    global _hip
    if _hip is None:
        from ._lib import _one_runtime
        _one_runtime()
        for name in ("libamdhip64.so", os.path.join("/opt/rocm/lib", "libamdhip64.so")):
            try:
                _hip = ctypes.CDLL(name)
                break
            except OSError:
                continue
        if _hip is None:
            raise ImportError("libamdhip64.so not found")
        _hip.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), ctypes.c_void_p, ctypes.c_void_p]
        _hip.hipEventRecord.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        _hip.hipEventSynchronize.argtypes = [ctypes.c_void_p]
        _hip.hipEventDestroy.argtypes = [ctypes.c_void_p]
        _hip.hipStreamSynchronize.argtypes = [ctypes.c_void_p]
        _hip.hipStreamDestroy.argtypes = [ctypes.c_void_p]
    return _hip


def _ok(r, what):
    if r != 0:
        raise RuntimeError("%s failed with HIP error %d" % (what, r))


def stream_create():
    s = ctypes.c_void_p()
    _ok(hip().hipStreamCreate(ctypes.byref(s)), "hipStreamCreate")
    return s.value


def stream_destroy(s):
    hip().hipStreamDestroy(s)


def stream_sync(s):
    _ok(hip().hipStreamSynchronize(s), "hipStreamSynchronize")


class Event:
    def __init__(self):
        e = ctypes.c_void_p()
        _ok(hip().hipEventCreate(ctypes.byref(e)), "hipEventCreate")
        self.h = e.value

    def record(self, stream):
        _ok(hip().hipEventRecord(self.h, stream), "hipEventRecord")

    def sync(self):
        _ok(hip().hipEventSynchronize(self.h), "hipEventSynchronize")

    def ms_since(self, start):
        ms = ctypes.c_float()
        _ok(hip().hipEventElapsedTime(ctypes.byref(ms), start.h, self.h), "hipEventElapsedTime")
        return ms.value

    def __del__(self):
        try:
            if self.h:
                hip().hipEventDestroy(self.h)
        except Exception:
            pass
