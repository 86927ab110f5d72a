"""Minimal ctypes binding of the HIP runtime libmptc_hip.so links against
(/opt/rocm/lib/libamdhip64.so.7): device buffers, streams and events for the
device-resident entry points.

One process must use one HIP runtime.  PyTorch wheels bundle their own copy of
the runtime; a process that drives libmptc_hip.so therefore allocates device
memory through this module rather than through torch.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
_hip = None

hipMemcpyHostToDevice, hipMemcpyDeviceToHost, hipMemcpyDeviceToDevice = 1, 2, 3


class HipError(RuntimeError):
    pass


def rt() -> C.CDLL:
    global _hip
    if _hip is None:
        _hip = C.CDLL(os.path.join(_ROCM, "lib", "libamdhip64.so.7"), mode=C.RTLD_GLOBAL)
        _hip.hipGetErrorString.restype = C.c_char_p
    return _hip


def check(rc: int, what: str = "hip"):
    if rc != 0:
        raise HipError(f"{what}: {rt().hipGetErrorString(rc).decode()} ({rc})")


def device_count() -> int:
    n = C.c_int(0)
    rc = rt().hipGetDeviceCount(C.byref(n))
    return n.value if rc == 0 else 0


def set_device(d: int = 0):
    check(rt().hipSetDevice(d), "hipSetDevice")


def synchronize():
    check(rt().hipDeviceSynchronize(), "hipDeviceSynchronize")


def device_name(d: int = 0) -> str:
    buf = C.create_string_buffer(256)
    if rt().hipDeviceGetName(buf, 256, d) == 0 and buf.value:
        return buf.value.decode()
    return f"HIP device {d}"  # the runtime may report no marketing name


class Buffer:
    """A device allocation of `nbytes` bytes."""

    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        p = C.c_void_p()
        check(rt().hipMalloc(C.byref(p), C.c_size_t(max(self.nbytes, 1))), "hipMalloc")
        self.ptr = p.value

    def free(self):
        if self.ptr:
            rt().hipFree(C.c_void_p(self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    def data_ptr(self) -> int:
        return self.ptr

    def offset(self, nbytes: int) -> "View":
        return View(self, nbytes)

    def upload(self, a: np.ndarray, offset: int = 0):
        a = np.ascontiguousarray(a)
        assert offset + a.nbytes <= self.nbytes
        check(rt().hipMemcpy(C.c_void_p(self.ptr + offset), a.ctypes.data_as(C.c_void_p),
                             C.c_size_t(a.nbytes), hipMemcpyHostToDevice), "hipMemcpy H2D")

    def download(self, dtype, count: int, offset: int = 0) -> np.ndarray:
        out = np.empty(count, dtype=dtype)
        assert offset + out.nbytes <= self.nbytes
        check(rt().hipMemcpy(out.ctypes.data_as(C.c_void_p), C.c_void_p(self.ptr + offset),
                             C.c_size_t(out.nbytes), hipMemcpyDeviceToHost), "hipMemcpy D2H")
        return out

    @staticmethod
    def from_array(a: np.ndarray) -> "Buffer":
        b = Buffer(a.nbytes)
        b.upload(a)
        return b


class View:
    """A pointer into a Buffer (keeps the buffer alive)."""

    def __init__(self, buf: Buffer, off: int):
        self.buf, self.off = buf, off

    def data_ptr(self) -> int:
        return self.buf.ptr + self.off


class HostBuffer:
    """Pinned (page-locked) host memory of `nbytes` bytes (hipHostMalloc), so
    that copies to and from the device run asynchronously at full PCIe rate."""

    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        p = C.c_void_p()
        check(rt().hipHostMalloc(C.byref(p), C.c_size_t(max(self.nbytes, 1)), C.c_uint(0)), "hipHostMalloc")
        self.ptr = p.value

    def array(self, dtype=np.uint8) -> np.ndarray:
        """A numpy view of the whole buffer."""
        n = self.nbytes // np.dtype(dtype).itemsize
        return np.ctypeslib.as_array(C.cast(C.c_void_p(self.ptr), C.POINTER(C.c_uint8)),
                                     shape=(self.nbytes,)).view(dtype)[:n]

    def data_ptr(self) -> int:
        return self.ptr

    def free(self):
        if self.ptr:
            rt().hipHostFree(C.c_void_p(self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def memcpy_async(dst: int, src: int, nbytes: int, kind: int, stream: "Stream"):
    """hipMemcpyAsync between raw addresses (kind: hipMemcpyHostToDevice, ...)."""
    check(rt().hipMemcpyAsync(C.c_void_p(dst), C.c_void_p(src), C.c_size_t(nbytes), C.c_int(kind),
                              C.c_void_p(stream.handle)), "hipMemcpyAsync")


class Stream:
    def __init__(self):
        s = C.c_void_p()
        check(rt().hipStreamCreate(C.byref(s)), "hipStreamCreate")
        self.handle = s.value

    def synchronize(self):
        check(rt().hipStreamSynchronize(C.c_void_p(self.handle)), "hipStreamSynchronize")

    def destroy(self):
        if self.handle:
            rt().hipStreamDestroy(C.c_void_p(self.handle))
            self.handle = None


class Event:
    def __init__(self):
        e = C.c_void_p()
        check(rt().hipEventCreate(C.byref(e)), "hipEventCreate")
        self.handle = e.value

    def record(self, stream: Stream | None = None):
        check(rt().hipEventRecord(C.c_void_p(self.handle),
                                  C.c_void_p(stream.handle if stream else None)), "hipEventRecord")

    def synchronize(self):
        check(rt().hipEventSynchronize(C.c_void_p(self.handle)), "hipEventSynchronize")

    def elapsed_ms(self, end: "Event") -> float:
        ms = C.c_float()
        check(rt().hipEventElapsedTime(C.byref(ms), C.c_void_p(self.handle), C.c_void_p(end.handle)),
              "hipEventElapsedTime")
        return float(ms.value)

    def destroy(self):
        if self.handle:
            rt().hipEventDestroy(C.c_void_p(self.handle))
            self.handle = None


def stream_wait(stream: Stream, event: Event):
    """Work queued on `stream` after this call waits for `event` (hipStreamWaitEvent)."""
    check(rt().hipStreamWaitEvent(C.c_void_p(stream.handle), C.c_void_p(event.handle), C.c_uint(0)),
          "hipStreamWaitEvent")
