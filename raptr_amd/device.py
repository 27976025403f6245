"""Device memory, streams and events over the C ABI (no torch in the loop)."""
import ctypes as C

import numpy as np

from . import _capi as capi


class DeviceArray:
    """A device allocation owned by Python, typed like a flat numpy array."""

    def __init__(self, count, dtype):
        self.dtype = np.dtype(dtype)
        self.count = int(count)
        self.nbytes = max(self.count * self.dtype.itemsize, 1)
        p = C.c_void_p()
        capi.call("raptr_malloc", C.byref(p), self.nbytes)
        self._ptr = p.value

    @property
    def ptr(self):
        return self._ptr

    @classmethod
    def from_numpy(cls, arr, dtype=None, stream=None):
        a = np.ascontiguousarray(arr, dtype=dtype)
        d = cls(a.size, a.dtype)
        if a.size:
            capi.call("raptr_memcpy_h2d", d.ptr, a.ctypes.data, a.nbytes, stream)
        return d

    @classmethod
    def zeros(cls, count, dtype, stream=None):
        d = cls(count, dtype)
        d.fill_zero(stream)
        return d

    def fill_zero(self, stream=None):
        capi.call("raptr_memset", self.ptr, 0, self.nbytes, stream)

    def numpy(self, count=None, stream=None):
        n = self.count if count is None else int(count)
        out = np.empty(n, self.dtype)
        if n:
            capi.call("raptr_memcpy_d2h", out.ctypes.data, self.ptr, n * self.dtype.itemsize,
                      stream)
        return out

    def copy_from(self, arr, stream=None):
        a = np.ascontiguousarray(arr, dtype=self.dtype)
        assert a.size <= self.count
        if a.size:
            capi.call("raptr_memcpy_h2d", self.ptr, a.ctypes.data, a.nbytes, stream)

    def offset(self, k):
        """Raw pointer to element k."""
        return self._ptr + int(k) * self.dtype.itemsize

    def free(self):
        if getattr(self, "_ptr", None):
            capi.lib().raptr_free(self._ptr)
            self._ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    def __len__(self):
        return self.count


class Workspace(DeviceArray):
    def __init__(self, nbytes):
        super().__init__(max(int(nbytes), 256), np.uint8)


class Stream:
    def __init__(self):
        s = C.c_void_p()
        capi.call("raptr_stream_create", C.byref(s))
        self.handle = s.value

    def sync(self):
        capi.call("raptr_stream_sync", self.handle)

    def wait(self, event):
        """Work queued on this stream from now on waits for `event`."""
        capi.call("raptr_stream_wait_event", self.handle, event.handle)

    def __del__(self):
        try:
            if self.handle:
                capi.lib().raptr_stream_destroy(self.handle)
                self.handle = None
        except Exception:
            pass


class Event:
    def __init__(self):
        e = C.c_void_p()
        capi.call("raptr_event_create", C.byref(e))
        self.handle = e.value

    def record(self, stream=None):
        capi.call("raptr_event_record", self.handle, stream)

    def elapsed_ms(self, end):
        ms = C.c_float()
        capi.call("raptr_event_elapsed_ms", C.byref(ms), self.handle, end.handle)
        return float(ms.value)

    def __del__(self):
        try:
            if self.handle:
                capi.lib().raptr_event_destroy(self.handle)
                self.handle = None
        except Exception:
            pass


def device_count():
    n = C.c_int()
    capi.call("raptr_device_count", C.byref(n))
    return n.value


def device_name():
    buf = C.create_string_buffer(256)
    capi.call("raptr_device_name", buf, 256)
    return buf.value.decode()


def synchronize():
    capi.call("raptr_device_sync")


def mem_info():
    f, t = C.c_size_t(), C.c_size_t()
    capi.call("raptr_mem_info", C.byref(f), C.byref(t))
    return f.value, t.value
