"""Device context and device arrays over the mxg C ABI (no torch needed)."""
from __future__ import annotations

import ctypes as C
import weakref

import numpy as np

from . import _native as N


class DeviceArray:
    """A device allocation holding a 1-D or 2-D (column-major) numpy-typed array."""

    def __init__(self, ctx: "Context", shape, dtype):
        self.ctx = ctx
        self.shape = tuple(int(s) for s in np.atleast_1d(shape))
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = C.c_void_p()
        N.check_gpu(N.gpu().mxg_malloc(ctx.handle, self.nbytes, C.byref(p)))
        self.ptr = p.value
        ctx.adopt(self, Context.ARRAYS)

    @property
    def size(self) -> int:
        return int(np.prod(self.shape))

    def upload(self, a: np.ndarray) -> "DeviceArray":
        a = np.asarray(a, dtype=self.dtype)
        assert a.size == self.size, (a.shape, self.shape)
        a = np.asfortranarray(a.reshape(self.shape, order="F")) if len(self.shape) > 1 else np.ascontiguousarray(a)
        N.check_gpu(N.gpu().mxg_memcpy_h2d(self.ctx.handle, self.ptr, a.ctypes.data, self.nbytes))
        return self

    def download(self) -> np.ndarray:
        out = np.empty(self.shape, dtype=self.dtype, order="F")
        N.check_gpu(N.gpu().mxg_memcpy_d2h(self.ctx.handle, out.ctypes.data, self.ptr, self.nbytes))
        return out

    def zero(self) -> "DeviceArray":
        N.check_gpu(N.gpu().mxg_memset(self.ctx.handle, self.ptr, 0, self.nbytes))
        return self

    def free(self) -> None:
        if self.ptr and self.ctx.handle:
            N.gpu().mxg_free(self.ctx.handle, self.ptr)
        self.ptr = None

    close = free

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Context:
    """One GPU: device selection and the stream every mxg call runs on.

    Objects built on the context (solvers, problems, device arrays) register
    here; close() releases the ones still alive, solvers first, before the
    context itself, so no native handle outlives the context it was made on."""

    SOLVERS, PROBLEMS, ARRAYS = 0, 1, 2

    def __init__(self, device: int = 0):
        h = C.c_void_p()
        N.check_gpu(N.gpu().mxg_init(device, C.byref(h)))
        self.handle = h.value
        self._owned = [weakref.WeakSet() for _ in range(3)]

    def adopt(self, obj, kind: int) -> None:
        self._owned[kind].add(obj)

    def array(self, a: np.ndarray, dtype=None) -> DeviceArray:
        a = np.asarray(a)
        d = DeviceArray(self, a.shape, dtype or a.dtype)
        return d.upload(a)

    def empty(self, shape, dtype=np.float64) -> DeviceArray:
        return DeviceArray(self, shape, dtype)

    def zeros(self, shape, dtype=np.float64) -> DeviceArray:
        return DeviceArray(self, shape, dtype).zero()

    def synchronize(self) -> None:
        N.check_gpu(N.gpu().mxg_synchronize(self.handle))

    def mem_info(self):
        """(free, total) bytes of device memory."""
        f, t = C.c_size_t(), C.c_size_t()
        N.check_gpu(N.gpu().mxg_mem_info(self.handle, C.byref(f), C.byref(t)))
        return f.value, t.value

    def close(self) -> None:
        if self.handle:
            for group in self._owned:
                for obj in list(group):
                    obj.close()
            N.gpu().mxg_finalize(self.handle)
            self.handle = None


def device_count() -> int:
    n = C.c_int(0)
    rc = N.gpu().mxg_device_count(C.byref(n))
    return n.value if rc == 0 else 0


def kernel_desc(name: str, diag: float = 1e3, alpha: float = 1.0) -> N.MxgKernel:
    return N.MxgKernel(N.KERNELS[name], float(diag), float(alpha))
