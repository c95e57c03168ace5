"""Device arrays: C-ordered buffers in HBM owned through the library's allocator.

A DeviceArray is a contiguous, C-ordered view (shape, dtype, byte offset) of a Storage block
allocated with gtn_malloc; reshapes and flat sub-ranges share the storage.  Host data moves in and out
only through gtn_memcpy.
"""
import ctypes as C

import numpy as np

from . import _lib as L


class Storage:
    __slots__ = ("ptr", "nbytes", "_lib")

    def __init__(self, nbytes):
        lib = L.lib()
        p = C.c_void_p()
        L.check(lib.gtn_malloc(int(nbytes), C.byref(p)), "gtn_malloc")
        self.ptr = p.value or 0
        self.nbytes = int(nbytes)
        self._lib = lib

    def __del__(self):
        try:
            if self.ptr:
                self._lib.gtn_free(C.c_void_p(self.ptr))
                self.ptr = 0
        except Exception:
            pass


def _numel(shape):
    n = 1
    for s in shape:
        n *= int(s)
    return n


class DeviceArray:
    __slots__ = ("storage", "offset", "shape", "dtype")

    def __init__(self, storage, offset, shape, dtype):
        self.storage = storage
        self.offset = int(offset)
        self.shape = tuple(int(s) for s in shape)
        self.dtype = np.dtype(dtype)

    # ---- construction -------------------------------------------------------------------------
    @staticmethod
    def empty(shape, dtype=np.float64):
        shape = tuple(int(s) for s in (shape if isinstance(shape, (tuple, list)) else (shape,)))
        dt = np.dtype(dtype)
        return DeviceArray(Storage(max(_numel(shape) * dt.itemsize, 1)), 0, shape, dt)

    @staticmethod
    def zeros(shape, dtype=np.float64):
        a = DeviceArray.empty(shape, dtype)
        a.fill_(0.0)
        return a

    @staticmethod
    def from_numpy(x):
        x = np.ascontiguousarray(x)
        a = DeviceArray.empty(x.shape, x.dtype)
        if x.size:
            L.check(L.lib().gtn_memcpy(C.c_void_p(a.data_ptr()), x.ctypes.data_as(C.c_void_p), x.nbytes, 1,
                                       L.stream_ptr()), "gtn_memcpy")
        return a

    # ---- properties ---------------------------------------------------------------------------
    def data_ptr(self):
        return self.storage.ptr + self.offset

    def numel(self):
        return _numel(self.shape)

    @property
    def ndim(self):
        return len(self.shape)

    def element_size(self):
        return self.dtype.itemsize

    @property
    def nbytes(self):
        return self.numel() * self.dtype.itemsize

    def is_contiguous(self):
        return True

    def is_complex(self):
        return self.dtype.kind == "c"

    def __len__(self):
        return self.shape[0]

    # ---- views and copies ---------------------------------------------------------------------
    def reshape(self, *shape):
        if len(shape) == 1 and isinstance(shape[0], (tuple, list)):
            shape = tuple(shape[0])
        shape = list(shape)
        if -1 in shape:
            k = shape.index(-1)
            rest = _numel([s for i, s in enumerate(shape) if i != k])
            shape[k] = self.numel() // max(rest, 1)
        if _numel(shape) != self.numel():
            raise ValueError("cannot reshape %s to %s" % (self.shape, tuple(shape)))
        return DeviceArray(self.storage, self.offset, shape, self.dtype)

    def flat(self, start, count, shape=None):
        """The elements [start, start + count) of the flattened array, as a view."""
        if start < 0 or start + count > self.numel():
            raise IndexError("flat range out of bounds")
        return DeviceArray(self.storage, self.offset + start * self.dtype.itemsize,
                           shape if shape is not None else (count,), self.dtype)

    def clone(self):
        out = DeviceArray.empty(self.shape, self.dtype)
        if self.numel():
            L.check(L.lib().gtn_memcpy(C.c_void_p(out.data_ptr()), C.c_void_p(self.data_ptr()), self.nbytes, 3,
                                       L.stream_ptr()), "gtn_memcpy")
        return out

    def copy_from(self, other):
        if other.shape != self.shape or other.dtype != self.dtype:
            raise ValueError("copy_from: shape or dtype mismatch")
        if self.numel():
            L.check(L.lib().gtn_memcpy(C.c_void_p(self.data_ptr()), C.c_void_p(other.data_ptr()), self.nbytes, 3,
                                       L.stream_ptr()), "gtn_memcpy")
        return self

    def fill_(self, value):
        if self.numel():
            L.check(L.lib().gtn_fill(C.c_void_p(self.data_ptr()), self.numel(), L.dtype_code(self), float(value),
                                     L.stream_ptr()), "gtn_fill")
        return self

    def numpy(self):
        out = np.empty(self.shape, dtype=self.dtype)
        if out.size:
            L.check(L.lib().gtn_memcpy(out.ctypes.data_as(C.c_void_p), C.c_void_p(self.data_ptr()), out.nbytes, 2,
                                       L.stream_ptr()), "gtn_memcpy")
        return out

    def cpu(self):
        return self

    def item(self):
        if self.numel() != 1:
            raise ValueError("item() needs a single element")
        return self.numpy().reshape(-1)[0].item()

    def astype_complex(self):
        """Complex copy of a real array (imaginary parts zero)."""
        h = self.numpy().astype(np.complex128)
        return DeviceArray.from_numpy(h)

    def __repr__(self):
        return "DeviceArray(shape=%s, dtype=%s)" % (self.shape, self.dtype)


def copy2d(dst, ldd, src, lds, rows, cols, dst_off=0, src_off=0):
    """dst[dst_off + r*ldd + c] = src[src_off + r*lds + c] for r < rows, c < cols (element offsets)."""
    if rows <= 0 or cols <= 0:
        return
    if dst_off + (rows - 1) * ldd + cols > dst.numel() or src_off + (rows - 1) * lds + cols > src.numel():
        raise IndexError("copy2d out of bounds")
    es = dst.element_size()
    L.check(L.lib().gtn_copy2d(C.c_void_p(dst.data_ptr() + es * dst_off), ldd, C.c_void_p(src.data_ptr() + es * src_off),
                               lds, rows, cols, L.dtype_code(dst), L.stream_ptr()), "gtn_copy2d")


def synchronize():
    L.check(L.lib().gtn_synchronize(), "gtn_synchronize")


class Event:
    """HIP event on the host layer's stream (for timing)."""

    def __init__(self):
        p = C.c_void_p()
        L.check(L.lib().gtn_event_create(C.byref(p)), "gtn_event_create")
        self.ptr = p

    def record(self):
        L.check(L.lib().gtn_event_record(self.ptr, L.stream_ptr()), "gtn_event_record")

    def elapsed_ms(self, stop):
        ms = C.c_float()
        L.check(L.lib().gtn_event_elapsed_ms(self.ptr, stop.ptr, C.byref(ms)), "gtn_event_elapsed_ms")
        return ms.value

    def __del__(self):
        try:
            L.lib().gtn_event_destroy(self.ptr)
        except Exception:
            pass
