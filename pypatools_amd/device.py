"""Device memory, streams and events on top of the C ABI.

DeviceArray owns one hipMalloc allocation; transfers are blocking on the
library's default stream unless a stream is given.  All particle and field
buffers of the engine are DeviceArrays: the engine never loads a second HIP
runtime (torch) into the process.
"""
import ctypes

import numpy as np

from . import _lib


def init(device=0):
    _lib.call("pp_init", device)


def device_count():
    c = ctypes.c_int(0)
    rc = _lib.lib().pp_device_count(ctypes.byref(c))
    return c.value if rc == 0 else 0


def mem_info():
    free, total = ctypes.c_size_t(), ctypes.c_size_t()
    _lib.call("pp_mem_info", ctypes.byref(free), ctypes.byref(total))
    return free.value, total.value


def synchronize(stream=None):
    _lib.call("pp_stream_sync", stream)


class DeviceArray:
    """A float64/uint8 device buffer with a NumPy-like shape."""

    __slots__ = ("ptr", "shape", "dtype", "nbytes", "__weakref__")

    def __init__(self, shape, dtype=np.float64):
        self.shape = tuple(int(s) for s in np.atleast_1d(shape)) if not isinstance(shape, tuple) else tuple(shape)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape, dtype=np.int64)) * self.dtype.itemsize
        p = ctypes.c_void_p()
        _lib.call("pp_malloc", ctypes.byref(p), max(self.nbytes, 1))
        self.ptr = p.value

    @property
    def size(self):
        return int(np.prod(self.shape, dtype=np.int64))

    @classmethod
    def from_host(cls, arr, dtype=None, stream=None):
        a = np.ascontiguousarray(arr, dtype=dtype if dtype is not None else np.asarray(arr).dtype)
        d = cls(a.shape, a.dtype)
        _lib.call("pp_memcpy_h2d", d.ptr, a.ctypes.data, a.nbytes, stream)
        return d

    def upload(self, arr, stream=None):
        a = np.ascontiguousarray(arr, dtype=self.dtype)
        if a.nbytes != self.nbytes:
            raise ValueError(f"size mismatch: {a.nbytes} vs {self.nbytes} bytes")
        _lib.call("pp_memcpy_h2d", self.ptr, a.ctypes.data, a.nbytes, stream)

    def to_host(self, out=None, stream=None):
        if out is None:
            out = np.empty(self.shape, self.dtype)
        if not out.flags.c_contiguous or out.nbytes != self.nbytes:
            raise ValueError("output must be C-contiguous with matching size")
        _lib.call("pp_memcpy_d2h", out.ctypes.data, self.ptr, self.nbytes, stream)
        return out

    def copy_from(self, other, stream=None):
        if other.nbytes != self.nbytes:
            raise ValueError("size mismatch")
        _lib.call("pp_memcpy_d2d", self.ptr, other.ptr, self.nbytes, stream)

    def fill(self, value, stream=None):
        if self.dtype != np.float64:
            raise TypeError("fill() is for float64 buffers")
        _lib.call("pp_fill_f64", self.ptr, float(value), self.size, stream)

    def view_ptr(self, offset_elems):
        return self.ptr + int(offset_elems) * self.dtype.itemsize

    def free(self):
        if self.ptr:
            try:
                _lib.lib().pp_free(self.ptr)
            except Exception:
                pass
            self.ptr = None

    def __del__(self):
        self.free()


class PinnedArray:
    """Page-locked host memory (pp_host_alloc) with a NumPy view, for asynchronous copies."""

    def __init__(self, nbytes):
        self.nbytes = max(int(nbytes), 1)
        p = ctypes.c_void_p()
        _lib.call("pp_host_alloc", ctypes.byref(p), self.nbytes)
        self.ptr = p.value
        self._buf = (ctypes.c_uint8 * self.nbytes).from_address(self.ptr)

    def view(self, dtype, count, offset_bytes=0):
        """A NumPy array of `count` items of `dtype` over this memory from `offset_bytes`."""
        dtype = np.dtype(dtype)
        if offset_bytes + count * dtype.itemsize > self.nbytes:
            raise ValueError("view beyond the pinned buffer")
        return np.frombuffer(self._buf, dtype=dtype, count=count, offset=offset_bytes)

    def free(self):
        if getattr(self, "ptr", None):
            self._buf = None
            try:
                _lib.lib().pp_host_free(self.ptr)
            except Exception:
                pass
            self.ptr = None

    def __del__(self):
        self.free()


def copy_h2d_async(dst_ptr, pinned, nbytes, stream=None, offset_bytes=0):
    _lib.call("pp_memcpy_h2d_async", dst_ptr, pinned.ptr + offset_bytes, nbytes, stream)


def copy_d2h_async(pinned, src_ptr, nbytes, stream=None, offset_bytes=0):
    _lib.call("pp_memcpy_d2h_async", pinned.ptr + offset_bytes, src_ptr, nbytes, stream)


class Stream:
    def __init__(self):
        h = ctypes.c_void_p()
        _lib.call("pp_stream_create", ctypes.byref(h))
        self.handle = h.value

    def sync(self):
        _lib.call("pp_stream_sync", self.handle)

    def __del__(self):
        if getattr(self, "handle", None):
            try:
                _lib.lib().pp_stream_destroy(self.handle)
            except Exception:
                pass
            self.handle = None


class Event:
    def __init__(self):
        h = ctypes.c_void_p()
        _lib.call("pp_event_create", ctypes.byref(h))
        self.handle = h.value

    def record(self, stream=None):
        _lib.call("pp_event_record", self.handle, stream)

    def elapsed_ms(self, stop):
        ms = ctypes.c_float()
        _lib.call("pp_event_elapsed_ms", ctypes.byref(ms), self.handle, stop.handle)
        return float(ms.value)

    def query(self):
        """True once the work recorded before this event has completed (does not block)."""
        done = ctypes.c_int32()
        _lib.call("pp_event_query", self.handle, ctypes.byref(done))
        return bool(done.value)

    def sync(self):
        _lib.call("pp_event_sync", self.handle)

    def __del__(self):
        if getattr(self, "handle", None):
            try:
                _lib.lib().pp_event_destroy(self.handle)
            except Exception:
                pass
            self.handle = None
