"""Device buffers, streams and events over the C ABI (pm_malloc / pm_memcpy /
pm_stream_* / pm_event_*), for the Python tests and the benchmark.  Only the
native library's own HIP runtime is ever loaded."""
import ctypes

import numpy as np

import proto_amd as pa


class DeviceArray:
    """A flat device allocation of n elements of a numpy dtype."""

    def __init__(self, n, dtype=np.float64):
        self.n = int(n)
        self.dtype = np.dtype(dtype)
        self.nbytes = self.n * self.dtype.itemsize
        p = ctypes.c_void_p()
        pa.call("pm_malloc", ctypes.byref(p), self.nbytes)
        self.ptr = p

    @classmethod
    def from_numpy(cls, a, stream=None):
        a = np.ascontiguousarray(a)
        d = cls(a.size, a.dtype)
        pa.call("pm_memcpy", d.ptr, a.ctypes.data_as(ctypes.c_void_p), d.nbytes, 0, stream)
        pa.call("pm_stream_sync", stream)
        return d

    @classmethod
    def zeros(cls, n, dtype=np.float64, stream=None):
        d = cls(n, dtype)
        pa.call("pm_memset_zero", d.ptr, d.nbytes, stream)
        return d

    def numpy(self, stream=None):
        out = np.empty(self.n, dtype=self.dtype)
        pa.call("pm_stream_sync", stream)
        pa.call("pm_memcpy", out.ctypes.data_as(ctypes.c_void_p), self.ptr, self.nbytes, 1, stream)
        pa.call("pm_stream_sync", stream)
        return out

    def upload(self, a, stream=None):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.size == self.n
        pa.call("pm_memcpy", self.ptr, a.ctypes.data_as(ctypes.c_void_p), self.nbytes, 0, stream)
        pa.call("pm_stream_sync", stream)

    def free(self):
        if self.ptr is not None and self.ptr.value:
            pa.lib().pm_free(self.ptr)
        self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    @property
    def vp(self):
        return self.ptr


class Stream:
    def __init__(self):
        h = ctypes.c_void_p()
        pa.call("pm_stream_create", ctypes.byref(h))
        self.h = h

    def sync(self):
        pa.call("pm_stream_sync", self.h)

    def __del__(self):
        try:
            if self.h:
                pa.lib().pm_stream_destroy(self.h)
                self.h = None
        except Exception:
            pass


class Event:
    def __init__(self):
        h = ctypes.c_void_p()
        pa.call("pm_event_create", ctypes.byref(h))
        self.h = h

    def record(self, stream=None):
        pa.call("pm_event_record", self.h, stream.h if isinstance(stream, Stream) else stream)

    def sync(self):
        pa.call("pm_event_sync", self.h)

    def elapsed_ms(self, end):
        ms = ctypes.c_float()
        pa.call("pm_event_elapsed_ms", self.h, end.h, ctypes.byref(ms))
        return ms.value

    def __del__(self):
        try:
            if self.h:
                pa.lib().pm_event_destroy(self.h)
                self.h = None
        except Exception:
            pass


def synchronize():
    pa.call("pm_device_sync")


def gpu_available():
    n = ctypes.c_int(0)
    pa.lib().pm_device_count(ctypes.byref(n))
    return n.value > 0
