"""Device memory, pinned host memory, streams and events through the library's
own HIP runtime (C ABI "device memory utilities" in include/rsh_scatter.h).

These replace torch as GPU plumbing: torch wheels bundle a different HIP
runtime, and two HIP runtimes cannot share one device inside a process.
"""
from __future__ import annotations

import ctypes as C
from typing import Sequence

import numpy as np

from . import _native as N

H2D, D2H, D2D = 0, 1, 2


def _stream_ptr(stream):
    if stream is None:
        return None
    return stream.ptr if isinstance(stream, Stream) else int(stream)


class DeviceArray:
    """Contiguous device array (C order)."""

    def __init__(self, shape, dtype=np.float64, device: int = 0):
        self.shape = tuple(int(s) for s in (shape if isinstance(shape, (tuple, list)) else (shape,)))
        self.dtype = np.dtype(dtype)
        self.device = device
        self.size = int(np.prod(self.shape)) if self.shape else 1
        self.nbytes = self.size * self.dtype.itemsize
        p = C.c_void_p()
        N.call("rsh_dev_alloc", device, self.nbytes, C.byref(p))
        self.ptr = p.value

    @classmethod
    def from_host(cls, a, device: int = 0) -> "DeviceArray":
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype, device)
        d.copy_from(a)
        return d

    def data_ptr(self) -> int:
        return self.ptr

    def numel(self) -> int:
        return self.size

    def copy_from(self, host, stream=None) -> None:
        host = np.ascontiguousarray(host, dtype=self.dtype)
        assert host.nbytes == self.nbytes, "size mismatch"
        N.call("rsh_dev_memcpy", self.ptr, host.ctypes.data, self.nbytes, H2D, _stream_ptr(stream))

    def to_host(self) -> np.ndarray:
        out = np.empty(self.shape, dtype=self.dtype)
        N.call("rsh_dev_memcpy", out.ctypes.data, self.ptr, self.nbytes, D2H, None)
        return out

    def fill(self, value: float, stream=None) -> None:
        assert self.dtype == np.float64
        N.call("rsh_dev_fill_f64", self.ptr, self.size, float(value), _stream_ptr(stream))

    def scale(self, factor: float, stream=None) -> None:
        assert self.dtype == np.float64
        N.call("rsh_dev_scale_f64", self.ptr, self.size, float(factor), _stream_ptr(stream))

    def copy_kernel_from(self, src: "DeviceArray", stream=None) -> None:
        """Device copy through the library's streaming copy kernel (not an SDMA copy)."""
        assert src.nbytes == self.nbytes and self.dtype == np.float64
        N.call("rsh_dev_copy_f64", self.ptr, src.ptr, self.size, _stream_ptr(stream))

    def gather_cells(self, cells: Sequence[int]) -> np.ndarray:
        """For a [M][N][F] array: host copy of [M][len(cells)][F] (one 2D copy per cell)."""
        M, Nc, F = self.shape
        item = self.dtype.itemsize
        out = np.empty((M, len(cells), F), dtype=self.dtype)
        for i, c in enumerate(cells):
            N.call("rsh_dev_memcpy2d", out.ctypes.data + i * F * item, len(cells) * F * item,
                   self.ptr + int(c) * F * item, Nc * F * item, F * item, M, D2H)
        return out

    def gather_columns(self, cols: Sequence[int]) -> np.ndarray:
        """For a [R][N] array: host copy of [R][len(cols)]."""
        R, Nc = self.shape
        item = self.dtype.itemsize
        out = np.empty((R, len(cols)), dtype=self.dtype)
        for i, c in enumerate(cols):
            N.call("rsh_dev_memcpy2d", out.ctypes.data + i * item, len(cols) * item, self.ptr + int(c) * item,
                   Nc * item, item, R, D2H)
        return out

    def free(self) -> None:
        if self.ptr:
            N.call("rsh_dev_free", self.ptr)
            self.ptr = 0

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class PinnedArray:
    """Page-locked host array (hipHostMalloc) with a numpy view in .a."""

    def __init__(self, shape, dtype=np.float64):
        shape = tuple(shape) if isinstance(shape, (tuple, list)) else (int(shape),)
        dtype = np.dtype(dtype)
        nbytes = int(np.prod(shape)) * dtype.itemsize
        p = C.c_void_p()
        N.call("rsh_host_alloc_pinned", nbytes, C.byref(p))
        self.ptr = p.value
        buf = (C.c_char * max(nbytes, 1)).from_address(self.ptr)
        self.a = np.frombuffer(buf, dtype=dtype, count=int(np.prod(shape))).reshape(shape)

    def data_ptr(self) -> int:
        return self.ptr

    def free(self) -> None:
        if self.ptr:
            self.a = None
            N.call("rsh_host_free_pinned", self.ptr)
            self.ptr = 0

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Stream:
    def __init__(self, device: int = 0):
        p = C.c_void_p()
        N.call("rsh_stream_create", device, C.byref(p))
        self.ptr = p.value

    @property
    def cuda_stream(self) -> int:
        return self.ptr

    def synchronize(self) -> None:
        N.call("rsh_stream_synchronize", self.ptr)

    def destroy(self) -> None:
        if self.ptr:
            N.call("rsh_stream_destroy", self.ptr)
            self.ptr = 0


class Event:
    def __init__(self):
        p = C.c_void_p()
        N.call("rsh_event_create", C.byref(p))
        self.ptr = p.value

    def record(self, stream=None) -> None:
        N.call("rsh_event_record", self.ptr, _stream_ptr(stream))

    def elapsed_ms(self, end: "Event") -> float:
        ms = C.c_float()
        N.call("rsh_event_elapsed_ms", self.ptr, end.ptr, C.byref(ms))
        return ms.value

    def destroy(self) -> None:
        if self.ptr:
            N.call("rsh_event_destroy", self.ptr)
            self.ptr = 0


def synchronize(device: int = 0) -> None:
    N.call("rsh_dev_synchronize", device)


def mem_info(device: int = 0):
    f, t = C.c_int64(), C.c_int64()
    N.call("rsh_dev_mem_info", device, C.byref(f), C.byref(t))
    return f.value, t.value
