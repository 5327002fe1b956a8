"""Device buffers, streams and events on top of the C-ABI.

The library allocates HBM itself (ch_malloc); numpy arrays are the host side.
"""
import ctypes
import threading

import numpy as np

from . import _lib


class Stream:
    """A HIP stream (default priority)."""

    def __init__(self, handle=None):
        self._own = handle is None
        if handle is None:
            h = ctypes.c_void_p()
            _lib.check(_lib.lib().ch_stream_create(ctypes.byref(h)))
            handle = h.value
        self.handle = handle

    def sync(self):
        _lib.check(_lib.lib().ch_stream_sync(self.handle))

    def wait(self, event):
        """Later work on this stream waits for `event`."""
        _lib.check(_lib.lib().ch_stream_wait(self.handle, event.handle))

    def close(self):
        if self._own and self.handle:
            _lib.lib().ch_stream_destroy(self.handle)
        self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


NULL_STREAM = None


def _h(stream):
    return stream.handle if isinstance(stream, Stream) else stream


def arm_launch_events(start, stop):
    """The calling thread's next gather or scatter launch records `start`
    and `stop` (Events) right around its kernel (ch_set_launch_events)."""
    _lib.check(_lib.lib().ch_set_launch_events(start.handle, stop.handle))


class Event:
    def __init__(self):
        h = ctypes.c_void_p()
        _lib.check(_lib.lib().ch_event_create(ctypes.byref(h)))
        self.handle = h.value

    def record(self, stream=None):
        _lib.check(_lib.lib().ch_event_record(self.handle, _h(stream)))
        return self

    def sync(self):
        _lib.check(_lib.lib().ch_event_sync(self.handle))

    def done(self):
        d = ctypes.c_int32()
        _lib.check(_lib.lib().ch_event_query(self.handle, ctypes.byref(d)))
        return bool(d.value)

    def elapsed_ms(self, later):
        ms = ctypes.c_float()
        _lib.check(_lib.lib().ch_event_elapsed(ctypes.byref(ms), self.handle,
                                               later.handle))
        return float(ms.value)

    def __del__(self):
        try:
            if self.handle:
                _lib.lib().ch_event_destroy(self.handle)
        except Exception:
            pass


class DeviceArray:
    """An HBM allocation with a numpy-like shape and dtype."""

    def __init__(self, shape, dtype=np.float32, zero=False, stream=None):
        self.shape = tuple(int(s) for s in np.atleast_1d(shape)) \
            if not isinstance(shape, tuple) else tuple(int(s) for s in shape)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape, dtype=np.int64)) * \
            self.dtype.itemsize
        p = ctypes.c_void_p()
        _lib.check(_lib.lib().ch_malloc(ctypes.byref(p), self.nbytes))
        self.ptr = p.value
        if zero:
            self.zero(stream)

    @classmethod
    def from_host(cls, arr, dtype=None, stream=None):
        a = np.ascontiguousarray(arr, dtype=dtype or np.asarray(arr).dtype)
        d = cls(a.shape, a.dtype)
        d.upload(a, stream)
        return d

    @property
    def size(self):
        return int(np.prod(self.shape, dtype=np.int64))

    def upload(self, arr, stream=None):
        a = np.ascontiguousarray(arr, dtype=self.dtype)
        if a.nbytes != self.nbytes:
            raise ValueError("upload of %d bytes into %d" %
                             (a.nbytes, self.nbytes))
        _lib.check(_lib.lib().ch_memcpy_h2d(self.ptr, a.ctypes.data,
                                            self.nbytes, _h(stream)))
        if stream is not None:
            # pageable source: make sure the copy has consumed it
            _lib.check(_lib.lib().ch_stream_sync(_h(stream)))
        return self

    def download(self, out=None, stream=None):
        if out is None:
            out = np.empty(self.shape, self.dtype)
        if not out.flags.c_contiguous or out.nbytes != self.nbytes or \
                out.dtype != self.dtype:
            raise ValueError("download target must be a contiguous %s %s" %
                             (self.dtype, self.shape))
        _lib.check(_lib.lib().ch_memcpy_d2h(out.ctypes.data, self.ptr,
                                            self.nbytes, _h(stream)))
        return out

    def download_async(self, host, offset=0, stream=None):
        """Queues a copy into the pinned HostBuffer `host` at byte `offset`;
        the data is there once an event recorded after it has completed."""
        if offset < 0 or offset + self.nbytes > host.nbytes:
            raise ValueError("download_async: %d bytes at %d into %d" %
                             (self.nbytes, offset, host.nbytes))
        _lib.check(_lib.lib().ch_memcpy_d2h_async(host.ptr + offset,
                                                  self.ptr, self.nbytes,
                                                  _h(stream)))

    def zero(self, stream=None):
        _lib.check(_lib.lib().ch_zero(self.ptr, self.nbytes, _h(stream)))
        return self

    def free(self):
        if self.ptr:
            _lib.lib().ch_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    def __repr__(self):
        return "DeviceArray(%s, %s, 0x%x)" % (self.shape, self.dtype,
                                             self.ptr or 0)


def mem_info():
    f = ctypes.c_size_t()
    t = ctypes.c_size_t()
    _lib.check(_lib.lib().ch_mem_info(ctypes.byref(f), ctypes.byref(t)))
    return f.value, t.value


def device_sync():
    _lib.check(_lib.lib().ch_device_sync())


class HostBuffer:
    """Pinned host memory (ch_host_alloc): the target of asynchronous
    device-to-host copies."""

    def __init__(self, nbytes):
        self.nbytes = int(nbytes)
        p = ctypes.c_void_p()
        _lib.check(_lib.lib().ch_host_alloc(ctypes.byref(p), self.nbytes))
        self.ptr = p.value

    def free(self):
        if self.ptr:
            _lib.lib().ch_host_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class _PinnedView:
    """Owner of a HostBuffer while numpy arrays view it; gives the buffer
    back to its pool when the last array is dropped."""

    def __init__(self, buf, pool, shape, dtype):
        self._buf = buf
        self._pool = pool
        dt = np.dtype(dtype)
        self.__array_interface__ = {
            "shape": tuple(shape), "typestr": dt.str,
            "data": (buf.ptr, False), "version": 3}

    def __del__(self):
        try:
            self._pool.put(self._buf)
        except Exception:
            pass


class HostPool:
    """Reusable pinned host buffers keyed by size.  as_array() hands a buffer
    out as a numpy array that owns it: the buffer returns to the pool when
    that array (and every view of it) is gone, so results are delivered
    without a host-side copy."""

    def __init__(self, max_free_bytes=4 << 30):
        self._free = {}
        self._bytes = 0
        self._max = max_free_bytes
        self._lock = threading.RLock()

    def get(self, nbytes):
        with self._lock:
            lst = self._free.get(int(nbytes))
            if lst:
                b = lst.pop()
                self._bytes -= b.nbytes
                return b
        return HostBuffer(nbytes)

    def put(self, buf):
        with self._lock:
            if self._bytes + buf.nbytes > self._max:
                buf.free()
                return
            self._free.setdefault(buf.nbytes, []).append(buf)
            self._bytes += buf.nbytes

    def as_array(self, buf, shape, dtype=np.float32):
        if int(np.prod(shape)) * np.dtype(dtype).itemsize > buf.nbytes:
            raise ValueError("as_array: %s %s does not fit %d bytes" %
                             (shape, dtype, buf.nbytes))
        return np.asarray(_PinnedView(buf, self, shape, dtype))

    def clear(self):
        with self._lock:
            for lst in self._free.values():
                for b in lst:
                    b.free()
            self._free.clear()
            self._bytes = 0
