"""Device memory, streams and events through the HIP runtime that libqtx.so uses.

A few ctypes calls into ROCm's libamdhip64 -- the same shared object libqtx.so
is linked against, so buffers, streams and events made here belong to the
runtime that runs the kernels.  (PyTorch ships its own copy of the HIP and HSA
runtimes; mixing the two in one process is avoided by not using it here.)
"""
import ctypes
import os

import numpy as np

_ROCM = os.environ.get('ROCM_PATH', '/opt/rocm')
_H2D, _D2H, _D2D = 1, 2, 3
_lib = None


def lib():
    global _lib
    if _lib is None:
        L = ctypes.CDLL(os.path.join(_ROCM, 'lib', 'libamdhip64.so.7'))
        vp, sz, i32 = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int
        L.hipGetErrorString.restype = ctypes.c_char_p
        L.hipGetErrorString.argtypes = [i32]
        L.hipMalloc.argtypes = [ctypes.POINTER(vp), sz]
        L.hipFree.argtypes = [vp]
        L.hipMemcpy.argtypes = [vp, vp, sz, i32]
        L.hipMemcpyAsync.argtypes = [vp, vp, sz, i32, vp]
        L.hipMemset.argtypes = [vp, i32, sz]
        L.hipMemsetAsync.argtypes = [vp, i32, sz, vp]
        L.hipHostMalloc.argtypes = [ctypes.POINTER(vp), sz, ctypes.c_uint]
        L.hipHostFree.argtypes = [vp]
        L.hipSetDevice.argtypes = [i32]
        L.hipGetDeviceCount.argtypes = [ctypes.POINTER(i32)]
        L.hipStreamCreateWithFlags.argtypes = [ctypes.POINTER(vp), ctypes.c_uint]
        L.hipStreamDestroy.argtypes = [vp]
        L.hipStreamSynchronize.argtypes = [vp]
        L.hipEventCreate.argtypes = [ctypes.POINTER(vp)]
        L.hipEventDestroy.argtypes = [vp]
        L.hipEventRecord.argtypes = [vp, vp]
        L.hipEventSynchronize.argtypes = [vp]
        L.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), vp, vp]
        _lib = L
    return _lib


class HipError(RuntimeError):
    pass


def check(err, what=''):
    if err != 0:
        raise HipError('%s: %s (%d)' % (what, lib().hipGetErrorString(err).decode(), err))


def device_count():
    n = ctypes.c_int()
    check(lib().hipGetDeviceCount(ctypes.byref(n)), 'hipGetDeviceCount')
    return n.value


def set_device(d):
    check(lib().hipSetDevice(d), 'hipSetDevice')


def synchronize():
    check(lib().hipDeviceSynchronize(), 'hipDeviceSynchronize')


class Buffer:
    """nbytes of device memory; .ptr is the device address."""

    def __init__(self, nbytes):
        self.nbytes = int(nbytes)
        p = ctypes.c_void_p()
        check(lib().hipMalloc(ctypes.byref(p), max(self.nbytes, 1)), 'hipMalloc(%d)' % self.nbytes)
        self.ptr = p.value

    @classmethod
    def from_array(cls, a):
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes)
        b.upload(a)
        return b

    def upload(self, a, stream=None):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        if stream is None:
            # a pageable-memory copy may return before its DMA lands: wait for the device
            check(lib().hipMemcpy(self.ptr, a.ctypes.data, a.nbytes, _H2D), 'hipMemcpy H2D')
            synchronize()
        else:
            check(lib().hipMemcpyAsync(self.ptr, a.ctypes.data, a.nbytes, _H2D, stream), 'hipMemcpyAsync H2D')

    def download(self, nbytes=None, offset=0):
        """Copy to a new host array; work on other streams must be finished first."""
        n = self.nbytes - offset if nbytes is None else int(nbytes)
        out = np.empty(n, np.uint8)
        check(lib().hipMemcpy(out.ctypes.data, self.ptr + offset, n, _D2H), 'hipMemcpy D2H')
        return out

    def zero(self, value=0):
        """Fill with value and wait (null-stream work does not order non-blocking streams)."""
        check(lib().hipMemset(self.ptr, value, self.nbytes), 'hipMemset')
        synchronize()

    def copy_from(self, other, nbytes=None):
        n = min(self.nbytes, other.nbytes) if nbytes is None else nbytes
        check(lib().hipMemcpy(self.ptr, other.ptr, n, _D2D), 'hipMemcpy D2D')
        synchronize()

    def free(self):
        if self.ptr:
            lib().hipFree(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Pinned:
    """nbytes of pinned (page-locked) host memory; .array is a uint8 numpy view."""

    def __init__(self, nbytes):
        self.nbytes = int(nbytes)
        p = ctypes.c_void_p()
        check(lib().hipHostMalloc(ctypes.byref(p), max(self.nbytes, 1), 0), 'hipHostMalloc(%d)' % self.nbytes)
        self.ptr = p.value
        self.array = np.ctypeslib.as_array((ctypes.c_uint8 * max(self.nbytes, 1)).from_address(self.ptr))[:self.nbytes]

    def free(self):
        if self.ptr:
            self.array = None
            lib().hipHostFree(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Stream:
    """A non-blocking stream at the default priority."""

    def __init__(self):
        s = ctypes.c_void_p()
        check(lib().hipStreamCreateWithFlags(ctypes.byref(s), 1), 'hipStreamCreateWithFlags')
        self.handle = s.value

    def synchronize(self):
        check(lib().hipStreamSynchronize(self.handle), 'hipStreamSynchronize')

    def destroy(self):
        if self.handle:
            lib().hipStreamDestroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


class Event:
    def __init__(self):
        e = ctypes.c_void_p()
        check(lib().hipEventCreate(ctypes.byref(e)), 'hipEventCreate')
        self.handle = e.value

    def record(self, stream=None):
        check(lib().hipEventRecord(self.handle, stream.handle if isinstance(stream, Stream) else stream),
              'hipEventRecord')

    def synchronize(self):
        check(lib().hipEventSynchronize(self.handle), 'hipEventSynchronize')

    def elapsed_ms(self, end):
        ms = ctypes.c_float()
        check(lib().hipEventElapsedTime(ctypes.byref(ms), self.handle, end.handle), 'hipEventElapsedTime')
        return ms.value

    def __del__(self):
        try:
            if self.handle:
                lib().hipEventDestroy(self.handle)
        except Exception:
            pass
