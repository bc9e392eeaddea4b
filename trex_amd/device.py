"""Device memory, streams and events over the C ABI (no torch dependency)."""

from __future__ import annotations

import ctypes

import numpy as np

from trex_amd import _lib


class Stream:
    def __init__(self):
        """A non-blocking stream (default priority)."""
        h = ctypes.c_void_p()
        _lib.call("trex_stream_create", ctypes.byref(h))
        self.handle = h.value

    def synchronize(self) -> None:
        _lib.call("trex_stream_sync", self.handle)

    def wait(self, event: "Event") -> None:
        """Later work on this stream waits for `event` (no host wait)."""
        _lib.call("trex_stream_wait_event", self.handle, event.handle)

    def close(self) -> None:
        if self.handle:
            _lib.lib().trex_stream_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Event:
    def __init__(self):
        h = ctypes.c_void_p()
        _lib.call("trex_event_create", ctypes.byref(h))
        self.handle = h.value

    def record(self, stream: Stream | None = None) -> None:
        _lib.call("trex_event_record", self.handle, stream.handle if stream else None)

    def query(self) -> bool:
        """True when all work recorded before the event has completed."""
        rc = _lib.lib().trex_event_query(self.handle)
        if rc < 0:
            _lib.check(rc, "trex_event_query")
        return rc == 0

    def synchronize(self) -> None:
        _lib.call("trex_event_sync", self.handle)

    def elapsed_ms(self, later: "Event") -> float:
        ms = ctypes.c_float()
        _lib.call("trex_event_elapsed_ms", self.handle, later.handle, ctypes.byref(ms))
        return ms.value

    def __del__(self):
        try:
            if self.handle:
                _lib.lib().trex_event_destroy(self.handle)
        except Exception:
            pass


class DeviceBuffer:
    """A block of device memory owned by this object."""

    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        h = ctypes.c_void_p()
        _lib.call("trex_malloc", ctypes.byref(h), max(self.nbytes, 1))
        self.ptr = h.value

    @classmethod
    def from_array(cls, arr: np.ndarray, stream: Stream | None = None) -> "DeviceBuffer":
        arr = np.ascontiguousarray(arr)
        buf = cls(arr.nbytes)
        buf.upload(arr, stream)
        if stream is None:
            _lib.call("trex_device_sync")
        return buf

    def upload(self, arr: np.ndarray, stream: Stream | None = None, offset: int = 0) -> None:
        arr = np.ascontiguousarray(arr)
        if offset + arr.nbytes > self.nbytes:
            raise ValueError("upload overruns the buffer")
        if arr.nbytes:
            _lib.call("trex_memcpy_h2d", self.ptr + offset, arr.ctypes.data, arr.nbytes,
                      stream.handle if stream else None)
            if stream is None:
                _lib.call("trex_device_sync")

    def download(self, dtype, count: int | None = None, stream: Stream | None = None, offset: int = 0) -> np.ndarray:
        dtype = np.dtype(dtype)
        if count is None:
            count = (self.nbytes - offset) // dtype.itemsize
        out = np.empty(count, dtype=dtype)
        if out.nbytes:
            _lib.call("trex_memcpy_d2h", out.ctypes.data, self.ptr + offset, out.nbytes,
                      stream.handle if stream else None)
        if stream is None:
            _lib.call("trex_device_sync")
        else:
            stream.synchronize()
        return out

    def free(self) -> None:
        if self.ptr:
            _lib.lib().trex_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def synchronize() -> None:
    _lib.call("trex_device_sync")


class PinnedBuffer:
    """Page-locked host memory (``trex_host_alloc``) seen as a uint8 numpy array,
    so host <-> device copies run asynchronously at full link speed."""

    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        h = ctypes.c_void_p()
        _lib.call("trex_host_alloc", ctypes.byref(h), max(self.nbytes, 1))
        self.ptr = h.value
        self.array = np.ctypeslib.as_array((ctypes.c_uint8 * max(self.nbytes, 1)).from_address(self.ptr))

    def free(self) -> None:
        if self.ptr:
            self.array = None
            _lib.lib().trex_host_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Graph:
    """Asynchronous launches recorded on a stream and replayed as one HIP graph
    (``trex_graph_*``)::

        with Graph.capture(stream) as g:
            engine.match_batch_device(batch, d_final, d_accept, stream)
        g.launch(stream)          # the same launches, same pointers

    Calls that copy from host memory, synchronise or allocate cannot be
    recorded; a failed capture is ended and dropped."""

    def __init__(self):
        self.handle = None

    @classmethod
    def capture(cls, stream: Stream) -> "_Capture":
        return _Capture(cls(), stream)

    def launch(self, stream: Stream | None = None) -> None:
        if not self.handle:
            raise RuntimeError("graph not captured")
        _lib.call("trex_graph_launch", self.handle, stream.handle if stream else None)

    def close(self) -> None:
        if self.handle:
            _lib.lib().trex_graph_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class _Capture:
    def __init__(self, graph: Graph, stream: Stream):
        self.graph, self.stream = graph, stream

    def __enter__(self) -> Graph:
        _lib.call("trex_graph_begin", self.stream.handle)
        return self.graph

    def __exit__(self, exc_type, exc, tb) -> bool:
        if exc_type is not None:
            _lib.lib().trex_graph_abort(self.stream.handle)
            return False
        h = ctypes.c_void_p()
        _lib.call("trex_graph_end", self.stream.handle, ctypes.byref(h))
        self.graph.handle = h.value
        return False
