"""ctypes binding of the product's C ABI (include/qtx.h) for the tests.

The parity tests call the library exactly as an FFI client would: through the
exported C entry points of quicktex_amd/libqtx.so.
"""
import ctypes
import os

import numpy as np

from ._cpu_libs import BC1Params

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(ROOT, 'quicktex_amd', 'libqtx.so')
HEADER = os.path.join(ROOT, 'include', 'qtx.h')

_U8P = ctypes.POINTER(ctypes.c_uint8)


def _ptr(a):
    return a.ctypes.data_as(_U8P)


class QtxError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__('qtx status %d: %s' % (status, msg))
        self.status = status


class Qtx:
    def __init__(self, path=LIB_PATH):
        self.lib = ctypes.CDLL(path)
        L = self.lib
        L.qtx_last_error.restype = ctypes.c_char_p
        L.qtx_ctx_create.argtypes = [ctypes.c_int32, ctypes.POINTER(ctypes.c_void_p)]
        L.qtx_ctx_destroy.argtypes = [ctypes.c_void_p]
        L.qtx_ctx_synchronize.argtypes = [ctypes.c_void_p]
        L.qtx_bc1_params_init.argtypes = [ctypes.POINTER(BC1Params), ctypes.c_uint32, ctypes.c_int32, ctypes.c_int32]
        L.qtx_bc1_params_clamp.argtypes = [ctypes.POINTER(BC1Params)]
        P = ctypes.POINTER(BC1Params)
        vp, i32, u32, sz = ctypes.c_void_p, ctypes.c_int32, ctypes.c_uint32, ctypes.c_size_t
        for n in ('qtx_bc1_encode', 'qtx_bc3_encode'):
            getattr(L, n).argtypes = [vp, P, _U8P, i32, i32, _U8P, sz]
        L.qtx_bc4_encode.argtypes = [vp, u32, _U8P, i32, i32, _U8P, sz]
        L.qtx_bc5_encode.argtypes = [vp, u32, u32, _U8P, i32, i32, _U8P, sz]
        L.qtx_bc1_decode.argtypes = [vp, i32, i32, _U8P, i32, i32, _U8P, sz]
        L.qtx_bc3_decode.argtypes = [vp, i32, _U8P, i32, i32, _U8P, sz]
        L.qtx_bc4_decode.argtypes = [vp, u32, _U8P, i32, i32, _U8P, sz]
        L.qtx_bc5_decode.argtypes = [vp, u32, u32, _U8P, i32, i32, _U8P, sz]
        for n in ('qtx_bc1_encode_device', 'qtx_bc3_encode_device'):
            getattr(L, n).argtypes = [vp, P, vp, i32, i32, vp, vp]
        L.qtx_bc4_encode_device.argtypes = [vp, u32, vp, i32, i32, vp, vp]
        L.qtx_bc5_encode_device.argtypes = [vp, u32, u32, vp, i32, i32, vp, vp]
        L.qtx_bc1_decode_device.argtypes = [vp, i32, i32, vp, i32, i32, vp, vp]
        L.qtx_bc3_decode_device.argtypes = [vp, i32, vp, i32, i32, vp, vp]
        L.qtx_bc4_decode_device.argtypes = [vp, u32, vp, i32, i32, vp, vp]
        L.qtx_bc5_decode_device.argtypes = [vp, u32, u32, vp, i32, i32, vp, vp]
        self._ctx = None

    def check(self, st):
        if st != 0:
            raise QtxError(st, self.lib.qtx_last_error().decode())

    @property
    def ctx(self):
        if self._ctx is None:
            c = ctypes.c_void_p()
            self.check(self.lib.qtx_ctx_create(0, ctypes.byref(c)))
            self._ctx = c
        return self._ctx

    def close(self):
        if self._ctx is not None:
            self.lib.qtx_ctx_destroy(self._ctx)
            self._ctx = None

    def params(self, level=5, color_mode=4, interpolator=0):
        p = BC1Params()
        self.check(self.lib.qtx_bc1_params_init(ctypes.byref(p), level, color_mode, interpolator))
        return p

    @staticmethod
    def _nb(w, h):
        return ((w + 3) // 4) * ((h + 3) // 4)

    def bc1_encode(self, rgba, params):
        h, w = rgba.shape[:2]
        rgba = np.ascontiguousarray(rgba, dtype=np.uint8)
        out = np.zeros(self._nb(w, h) * 8, np.uint8)
        self.check(self.lib.qtx_bc1_encode(self.ctx, ctypes.byref(params), _ptr(rgba), w, h, _ptr(out), out.size))
        return out

    def bc3_encode(self, rgba, params):
        h, w = rgba.shape[:2]
        rgba = np.ascontiguousarray(rgba, dtype=np.uint8)
        out = np.zeros(self._nb(w, h) * 16, np.uint8)
        self.check(self.lib.qtx_bc3_encode(self.ctx, ctypes.byref(params), _ptr(rgba), w, h, _ptr(out), out.size))
        return out

    def bc4_encode(self, rgba, channel=3):
        h, w = rgba.shape[:2]
        rgba = np.ascontiguousarray(rgba, dtype=np.uint8)
        out = np.zeros(self._nb(w, h) * 8, np.uint8)
        self.check(self.lib.qtx_bc4_encode(self.ctx, channel, _ptr(rgba), w, h, _ptr(out), out.size))
        return out

    def bc5_encode(self, rgba, chan0=0, chan1=1):
        h, w = rgba.shape[:2]
        rgba = np.ascontiguousarray(rgba, dtype=np.uint8)
        out = np.zeros(self._nb(w, h) * 16, np.uint8)
        self.check(self.lib.qtx_bc5_encode(self.ctx, chan0, chan1, _ptr(rgba), w, h, _ptr(out), out.size))
        return out

    def bc1_decode(self, blocks, w, h, interpolator=0, write_alpha=False):
        blocks = np.ascontiguousarray(blocks, dtype=np.uint8)
        out = np.zeros((h, w, 4), np.uint8)
        self.check(self.lib.qtx_bc1_decode(self.ctx, interpolator, int(write_alpha), _ptr(blocks), w, h, _ptr(out), out.size))
        return out

    def bc3_decode(self, blocks, w, h, interpolator=0):
        blocks = np.ascontiguousarray(blocks, dtype=np.uint8)
        out = np.zeros((h, w, 4), np.uint8)
        self.check(self.lib.qtx_bc3_decode(self.ctx, interpolator, _ptr(blocks), w, h, _ptr(out), out.size))
        return out

    def bc4_decode(self, blocks, w, h, channel=3):
        blocks = np.ascontiguousarray(blocks, dtype=np.uint8)
        out = np.zeros((h, w, 4), np.uint8)
        self.check(self.lib.qtx_bc4_decode(self.ctx, channel, _ptr(blocks), w, h, _ptr(out), out.size))
        return out

    def bc5_decode(self, blocks, w, h, chan0=0, chan1=1):
        blocks = np.ascontiguousarray(blocks, dtype=np.uint8)
        out = np.zeros((h, w, 4), np.uint8)
        self.check(self.lib.qtx_bc5_decode(self.ctx, chan0, chan1, _ptr(blocks), w, h, _ptr(out), out.size))
        return out


def declared_symbols(header=HEADER):
    """Every function name declared in include/qtx.h."""
    import re
    src = open(header).read()
    return sorted(set(re.findall(r'\b(qtx_[a-z0-9_]+)\s*\(', src)))
