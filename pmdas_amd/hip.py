"""Minimal ctypes access to the HIP runtime (device buffers, copies, streams,
events) for the Python mirror, tests and the benchmark.  Plumbing only."""
import ctypes
import os

import numpy as np

_hip = None

H2D, D2H, D2D = 1, 2, 3


def lib():
    global _hip
    if _hip is None:
        for cand in ("libamdhip64.so", "/opt/rocm/lib/libamdhip64.so"):
            try:
                _hip = ctypes.CDLL(cand)
                break
            except OSError:
                continue
        if _hip is None:
            raise ImportError("libamdhip64.so not found")
        _hip.hipGetErrorString.restype = ctypes.c_char_p
    return _hip


def _chk(rc, what):
    if rc != 0:
        msg = lib().hipGetErrorString(rc).decode()
        raise RuntimeError(f"{what} failed: {msg} ({rc})")


def device_count():
    n = ctypes.c_int(0)
    rc = lib().hipGetDeviceCount(ctypes.byref(n))
    return n.value if rc == 0 else 0


def synchronize():
    _chk(lib().hipDeviceSynchronize(), "hipDeviceSynchronize")


def mem_info():
    """(free, total) bytes of device memory."""
    free, total = ctypes.c_size_t(0), ctypes.c_size_t(0)
    _chk(lib().hipMemGetInfo(ctypes.byref(free), ctypes.byref(total)), "hipMemGetInfo")
    return free.value, total.value


class DeviceBuffer:
    """An owned hipMalloc allocation."""

    def __init__(self, nbytes):
        self.nbytes = int(nbytes)
        p = ctypes.c_void_p()
        _chk(lib().hipMalloc(ctypes.byref(p), ctypes.c_size_t(max(self.nbytes, 1))), "hipMalloc")
        self.ptr = p.value

    def upload(self, arr):
        arr = np.ascontiguousarray(arr)
        assert arr.nbytes <= self.nbytes
        _chk(lib().hipMemcpy(ctypes.c_void_p(self.ptr), arr.ctypes.data_as(ctypes.c_void_p),
                             ctypes.c_size_t(arr.nbytes), H2D), "hipMemcpy H2D")

    def download(self, dtype, count):
        out = np.empty(count, dtype=dtype)
        _chk(lib().hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(self.ptr),
                             ctypes.c_size_t(out.nbytes), D2H), "hipMemcpy D2H")
        return out

    def free(self):
        if self.ptr:
            lib().hipFree(ctypes.c_void_p(self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def download(ptr, dtype, count):
    out = np.empty(count, dtype=dtype)
    _chk(lib().hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(ptr),
                         ctypes.c_size_t(out.nbytes), D2H), "hipMemcpy D2H")
    return out


class Stream:
    def __init__(self):
        s = ctypes.c_void_p()
        _chk(lib().hipStreamCreate(ctypes.byref(s)), "hipStreamCreate")
        self.handle = s.value

    def synchronize(self):
        _chk(lib().hipStreamSynchronize(ctypes.c_void_p(self.handle)), "hipStreamSynchronize")

    def destroy(self):
        if self.handle:
            lib().hipStreamDestroy(ctypes.c_void_p(self.handle))
            self.handle = None


class Event:
    def __init__(self):
        e = ctypes.c_void_p()
        _chk(lib().hipEventCreate(ctypes.byref(e)), "hipEventCreate")
        self.handle = e.value

    def record(self, stream=None):
        _chk(lib().hipEventRecord(ctypes.c_void_p(self.handle),
                                  ctypes.c_void_p(stream.handle if stream else None)),
             "hipEventRecord")

    def synchronize(self):
        _chk(lib().hipEventSynchronize(ctypes.c_void_p(self.handle)), "hipEventSynchronize")

    def elapsed_ms(self, end):
        ms = ctypes.c_float()
        _chk(lib().hipEventElapsedTime(ctypes.byref(ms), ctypes.c_void_p(self.handle),
                                       ctypes.c_void_p(end.handle)), "hipEventElapsedTime")
        return ms.value

    def destroy(self):
        if self.handle:
            lib().hipEventDestroy(ctypes.c_void_p(self.handle))
            self.handle = None


def device_name():
    buf = ctypes.create_string_buffer(256)
    if lib().hipDeviceGetName(buf, ctypes.c_int(256), ctypes.c_int(0)) != 0:
        return "unknown"
    return buf.value.decode(errors="replace")


def visible():
    return os.environ.get("HIP_VISIBLE_DEVICES", "")


class Graph:
    """Stream capture into a hipGraph (begin / end / replay) for callers that
    enqueue library calls into a graph."""

    def __init__(self, stream):
        self.stream = stream
        self.graph = None
        self.exec = None

    def begin(self):
        # mode 1: hipStreamCaptureModeThreadLocal
        _chk(lib().hipStreamBeginCapture(ctypes.c_void_p(self.stream.handle), ctypes.c_int(1)),
             "hipStreamBeginCapture")

    def end(self):
        g = ctypes.c_void_p()
        _chk(lib().hipStreamEndCapture(ctypes.c_void_p(self.stream.handle), ctypes.byref(g)),
             "hipStreamEndCapture")
        self.graph = g.value
        e = ctypes.c_void_p()
        _chk(lib().hipGraphInstantiate(ctypes.byref(e), ctypes.c_void_p(self.graph), None, None,
                                       ctypes.c_size_t(0)), "hipGraphInstantiate")
        self.exec = e.value

    def launch(self):
        _chk(lib().hipGraphLaunch(ctypes.c_void_p(self.exec), ctypes.c_void_p(self.stream.handle)),
             "hipGraphLaunch")

    def destroy(self):
        if self.exec:
            lib().hipGraphExecDestroy(ctypes.c_void_p(self.exec))
            self.exec = None
        if self.graph:
            lib().hipGraphDestroy(ctypes.c_void_p(self.graph))
            self.graph = None
