"""ctypes binding of the C ABI in include/qtx.h (quicktex_amd/libqtx.so).

A thin client for callers that hold their own device memory (torch tensors,
hipMalloc'ed buffers) and want the ``*_device`` entry points, and for FFI-style
tests.  The drop-in Python API is the pybind11 module (quicktex_amd.s3tc.*).

Raises ``RuntimeError`` when libqtx.so has not been built: there is no CPU path.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, 'libqtx.so')

INTERP_IDEAL, INTERP_IDEAL_ROUND, INTERP_NVIDIA, INTERP_AMD = 0, 1, 2, 3
COLOR_FOUR, COLOR_THREE, COLOR_THREE_BLACK = 4, 3, 0x13
FMT_BC1, FMT_BC3, FMT_BC4, FMT_BC5 = 1, 3, 4, 5
BC1_AUTO, BC1_LANE, BC1_GROUP = 0, 1, 2


class BC1Params(ctypes.Structure):
    """qtx_bc1_params (include/qtx.h)."""
    _fields_ = [(n, ctypes.c_int32) for n in (
        'interpolator', 'color_mode', 'error_mode', 'endpoint_mode',
        'two_ls_passes', 'two_ep_passes', 'two_cf_passes', 'exhaustive')] + \
        [(n, ctypes.c_uint32) for n in ('power_iterations', 'search_rounds', 'orderings4', 'orderings3')]

    def as_tuple(self):
        return tuple(getattr(self, n) for n, _ in self._fields_)


OP_ENCODE, OP_DECODE = 0, 1


class Job(ctypes.Structure):
    """qtx_job (include/qtx.h): one texture of a batch."""
    _fields_ = [('op', ctypes.c_int32), ('format', ctypes.c_int32), ('width', ctypes.c_int32),
                ('height', ctypes.c_int32), ('src', ctypes.c_void_p), ('dst', ctypes.c_void_p),
                ('bc1', BC1Params), ('chan0', ctypes.c_int32), ('chan1', ctypes.c_int32),
                ('write_alpha', ctypes.c_int32), ('reserved', ctypes.c_int32)]


class BatchTuning(ctypes.Structure):
    """qtx_batch_tuning (include/qtx.h)."""
    _fields_ = [('workgroups_per_cu', ctypes.c_int32), ('unit_kops', ctypes.c_uint32)]


class ServerConfig(ctypes.Structure):
    """qtx_server_config (include/qtx.h); zero fields take the defaults."""
    _fields_ = [('device_bytes', ctypes.c_size_t), ('staging_bytes', ctypes.c_size_t), ('slots', ctypes.c_int32),
                ('max_batch_jobs', ctypes.c_int32), ('max_queue', ctypes.c_int32), ('unit_kops', ctypes.c_uint32),
                ('batch_bytes', ctypes.c_size_t)]


class ServerResult(ctypes.Structure):
    """qtx_server_result (include/qtx.h)."""
    _fields_ = [('status', ctypes.c_int32), ('batch_jobs', ctypes.c_int32), ('submit_s', ctypes.c_double),
                ('admit_s', ctypes.c_double), ('launch_s', ctypes.c_double), ('done_s', ctypes.c_double),
                ('message', ctypes.c_char * 160)]


class ServerStats(ctypes.Structure):
    """qtx_server_stats (include/qtx.h)."""
    _fields_ = [(n, ctypes.c_uint64) for n in ('submitted', 'completed', 'failed', 'batches', 'pixels',
                                                'max_waiting', 'max_batch_jobs')] + \
        [('device_bytes_peak', ctypes.c_size_t), ('busy_s', ctypes.c_double), ('cancelled', ctypes.c_uint64)]


def job(op, fmt, w, h, src, dst, params=None, chan0=0, chan1=1, write_alpha=False, interpolator=None):
    """A Job; params is a BC1Params for BC1/BC3 encodes, interpolator for BC1/BC3 decodes."""
    j = Job()
    j.op, j.format, j.width, j.height = op, fmt, w, h
    j.src, j.dst = src, dst
    if params is not None:
        j.bc1 = params
    if interpolator is not None:
        j.bc1.interpolator = interpolator
    j.chan0, j.chan1, j.write_alpha = chan0, chan1, int(write_alpha)
    return j


class QtxError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__('qtx status %d: %s' % (status, msg))
        self.status = status


_U8P = ctypes.POINTER(ctypes.c_uint8)


def _ptr(a):
    return a.ctypes.data_as(_U8P)


def _nb(w, h):
    return ((w + 3) // 4) * ((h + 3) // 4)


class Qtx:
    """One library handle and one lazily created context on ``device``."""

    def __init__(self, path=LIB_PATH, device=0):
        if not os.path.exists(path):
            raise RuntimeError('%s is not built; run `python -m quicktex_amd.build`' % path)
        self.lib = ctypes.CDLL(path)
        self.device = device
        L = self.lib
        vp, i32, u32, sz = ctypes.c_void_p, ctypes.c_int32, ctypes.c_uint32, ctypes.c_size_t
        P = ctypes.POINTER(BC1Params)
        L.qtx_last_error.restype = ctypes.c_char_p
        L.qtx_ctx_create.argtypes = [i32, ctypes.POINTER(vp)]
        L.qtx_ctx_destroy.argtypes = [vp]
        L.qtx_ctx_synchronize.argtypes = [vp]
        L.qtx_bc1_params_init.argtypes = [P, u32, i32, i32]
        L.qtx_bc1_params_clamp.argtypes = [P]
        L.qtx_ctx_set_bc1_encoder.argtypes = [vp, i32]
        for n in ('qtx_bc1_encode', 'qtx_bc3_encode'):
            getattr(L, n).argtypes = [vp, P, _U8P, i32, i32, _U8P, sz]
        L.qtx_bc4_encode.argtypes = [vp, u32, _U8P, i32, i32, _U8P, sz]
        L.qtx_bc5_encode.argtypes = [vp, u32, u32, _U8P, i32, i32, _U8P, sz]
        L.qtx_bc1_decode.argtypes = [vp, i32, i32, _U8P, i32, i32, _U8P, sz]
        L.qtx_bc3_decode.argtypes = [vp, i32, _U8P, i32, i32, _U8P, sz]
        L.qtx_bc4_decode.argtypes = [vp, u32, _U8P, i32, i32, _U8P, sz]
        L.qtx_bc5_decode.argtypes = [vp, u32, u32, _U8P, i32, i32, _U8P, sz]
        for n in ('qtx_bc1_encode_device', 'qtx_bc3_encode_device', 'qtx_bc1_encode_device_noalloc'):
            getattr(L, n).argtypes = [vp, P, vp, i32, i32, vp, vp]
        L.qtx_bc4_encode_device.argtypes = [vp, u32, vp, i32, i32, vp, vp]
        L.qtx_bc5_encode_device.argtypes = [vp, u32, u32, vp, i32, i32, vp, vp]
        L.qtx_bc1_decode_device.argtypes = [vp, i32, i32, vp, i32, i32, vp, vp]
        L.qtx_bc3_decode_device.argtypes = [vp, i32, vp, i32, i32, vp, vp]
        L.qtx_bc4_decode_device.argtypes = [vp, u32, vp, i32, i32, vp, vp]
        L.qtx_bc5_decode_device.argtypes = [vp, u32, u32, vp, i32, i32, vp, vp]
        JP = ctypes.POINTER(Job)
        TP = ctypes.POINTER(BatchTuning)
        L.qtx_batch_create.argtypes = [vp, JP, sz, ctypes.POINTER(vp)]
        L.qtx_batch_launch.argtypes = [vp, vp]
        L.qtx_batch_autotune.argtypes = [vp, vp, i32, TP, ctypes.POINTER(ctypes.c_float)]
        L.qtx_batch_set_tuning.argtypes = [vp, TP]
        L.qtx_batch_get_tuning.argtypes = [vp, TP]
        L.qtx_batch_destroy.argtypes = [vp]
        L.qtx_batch_run.argtypes = [vp, JP, sz, vp]
        L.qtx_batch_run_host.argtypes = [vp, JP, ctypes.POINTER(sz), sz]
        L.qtx_server_start.argtypes = [vp, ctypes.POINTER(ServerConfig), ctypes.POINTER(vp)]
        L.qtx_server_submit.argtypes = [vp, JP, sz, ctypes.POINTER(ctypes.c_uint64)]
        L.qtx_server_wait.argtypes = [vp, ctypes.c_uint64, i32, ctypes.POINTER(ServerResult)]
        L.qtx_server_stats_get.argtypes = [vp, ctypes.POINTER(ServerStats)]
        L.qtx_server_cancel.argtypes = [vp, ctypes.c_uint64]
        L.qtx_server_now.argtypes = [vp]
        L.qtx_server_now.restype = ctypes.c_double
        L.qtx_server_stop.argtypes = [vp]
        PP = ctypes.POINTER(ctypes.c_void_p)
        L.qtx_mip_count.argtypes = [i32, i32, i32]
        L.qtx_mip_count.restype = i32
        L.qtx_mips_workspace_bytes.argtypes = [i32, i32, i32]
        L.qtx_mips_workspace_bytes.restype = sz
        L.qtx_mips_device.argtypes = [vp, vp, i32, i32, i32, PP, vp, sz, vp]
        L.qtx_mips_host.argtypes = [vp, _U8P, i32, i32, i32, PP, ctypes.POINTER(sz)]
        L.qtx_mips_encode_host.argtypes = [vp, JP, _U8P, i32, i32, i32, PP, ctypes.POINTER(sz)]
        I32P = ctypes.POINTER(i32)
        L.qtx_has_alpha_device.argtypes = [vp, vp, i32, i32, vp, I32P]
        L.qtx_has_alpha_host.argtypes = [vp, _U8P, i32, i32, I32P]
        L.qtx_mips_encode_auto_host.argtypes = [vp, P, _U8P, i32, i32, i32, PP, ctypes.POINTER(sz), I32P]
        L.qtx_swizzle_bc3nm_device.argtypes = [vp, vp, i32, i32, vp]
        L.qtx_swizzle_bc3nm_host.argtypes = [vp, _U8P, i32, i32, _U8P]
        self._ctx = None

    # -- plumbing -------------------------------------------------------------
    def check(self, st):
        if st != 0:
            raise QtxError(st, self.lib.qtx_last_error().decode())

    @property
    def ctx(self):
        if self._ctx is None:
            c = ctypes.c_void_p()
            self.check(self.lib.qtx_ctx_create(self.device, ctypes.byref(c)))
            self._ctx = c
        return self._ctx

    def close(self):
        if self._ctx is not None:
            self.lib.qtx_ctx_destroy(self._ctx)
            self._ctx = None

    def set_bc1_encoder(self, which):
        """BC1_AUTO, BC1_LANE or BC1_GROUP for this context's later BC1/BC3 encodes."""
        self.check(self.lib.qtx_ctx_set_bc1_encoder(self.ctx, which))

    def synchronize(self):
        self.check(self.lib.qtx_ctx_synchronize(self.ctx))

    def params(self, level=5, color_mode=COLOR_FOUR, interpolator=INTERP_IDEAL):
        p = BC1Params()
        self.check(self.lib.qtx_bc1_params_init(ctypes.byref(p), level, color_mode, interpolator))
        return p

    # -- host buffers (numpy) ---------------------------------------------------
    def bc1_encode(self, rgba, params):
        h, w = rgba.shape[:2]
        rgba = np.ascontiguousarray(rgba, dtype=np.uint8)
        out = np.zeros(_nb(w, h) * 8, np.uint8)
        self.check(self.lib.qtx_bc1_encode(self.ctx, ctypes.byref(params), _ptr(rgba), w, h, _ptr(out), out.size))
        return out

    def bc3_encode(self, rgba, params):
        h, w = rgba.shape[:2]
        rgba = np.ascontiguousarray(rgba, dtype=np.uint8)
        out = np.zeros(_nb(w, h) * 16, np.uint8)
        self.check(self.lib.qtx_bc3_encode(self.ctx, ctypes.byref(params), _ptr(rgba), w, h, _ptr(out), out.size))
        return out

    def bc4_encode(self, rgba, channel=3):
        h, w = rgba.shape[:2]
        rgba = np.ascontiguousarray(rgba, dtype=np.uint8)
        out = np.zeros(_nb(w, h) * 8, np.uint8)
        self.check(self.lib.qtx_bc4_encode(self.ctx, channel, _ptr(rgba), w, h, _ptr(out), out.size))
        return out

    def bc5_encode(self, rgba, chan0=0, chan1=1):
        h, w = rgba.shape[:2]
        rgba = np.ascontiguousarray(rgba, dtype=np.uint8)
        out = np.zeros(_nb(w, h) * 16, np.uint8)
        self.check(self.lib.qtx_bc5_encode(self.ctx, chan0, chan1, _ptr(rgba), w, h, _ptr(out), out.size))
        return out

    def bc1_decode(self, blocks, w, h, interpolator=INTERP_IDEAL, write_alpha=False):
        blocks = np.ascontiguousarray(blocks, dtype=np.uint8)
        out = np.zeros((h, w, 4), np.uint8)
        self.check(self.lib.qtx_bc1_decode(self.ctx, interpolator, int(write_alpha), _ptr(blocks), w, h, _ptr(out),
                                           out.size))
        return out

    def bc3_decode(self, blocks, w, h, interpolator=INTERP_IDEAL):
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

    # -- device buffers (raw pointers, e.g. tensor.data_ptr()) ------------------
    def bc1_encode_device(self, params, d_rgba, w, h, d_blocks, stream=0):
        self.check(self.lib.qtx_bc1_encode_device(self.ctx, ctypes.byref(params), d_rgba, w, h, d_blocks, stream))

    def bc1_encode_device_noalloc(self, params, d_rgba, w, h, d_blocks, stream=0):
        """qtx_bc1_encode_device without allocation (the library's static scratch)."""
        self.check(self.lib.qtx_bc1_encode_device_noalloc(self.ctx, ctypes.byref(params), d_rgba, w, h, d_blocks,
                                                          stream))

    def bc3_encode_device(self, params, d_rgba, w, h, d_blocks, stream=0):
        self.check(self.lib.qtx_bc3_encode_device(self.ctx, ctypes.byref(params), d_rgba, w, h, d_blocks, stream))

    def bc4_encode_device(self, channel, d_rgba, w, h, d_blocks, stream=0):
        self.check(self.lib.qtx_bc4_encode_device(self.ctx, channel, d_rgba, w, h, d_blocks, stream))

    def bc5_encode_device(self, chan0, chan1, d_rgba, w, h, d_blocks, stream=0):
        self.check(self.lib.qtx_bc5_encode_device(self.ctx, chan0, chan1, d_rgba, w, h, d_blocks, stream))

    def bc1_decode_device(self, interpolator, write_alpha, d_blocks, w, h, d_rgba, stream=0):
        self.check(self.lib.qtx_bc1_decode_device(self.ctx, interpolator, int(write_alpha), d_blocks, w, h, d_rgba, stream))

    def bc3_decode_device(self, interpolator, d_blocks, w, h, d_rgba, stream=0):
        self.check(self.lib.qtx_bc3_decode_device(self.ctx, interpolator, d_blocks, w, h, d_rgba, stream))

    def bc4_decode_device(self, channel, d_blocks, w, h, d_rgba, stream=0):
        self.check(self.lib.qtx_bc4_decode_device(self.ctx, channel, d_blocks, w, h, d_rgba, stream))

    def bc5_decode_device(self, chan0, chan1, d_blocks, w, h, d_rgba, stream=0):
        self.check(self.lib.qtx_bc5_decode_device(self.ctx, chan0, chan1, d_blocks, w, h, d_rgba, stream))


class Batch:
    """A qtx_batch over device buffers: create once, launch many times."""

    def __init__(self, q, jobs):
        self.q = q
        arr = (Job * len(jobs))(*jobs)
        h = ctypes.c_void_p()
        q.check(q.lib.qtx_batch_create(q.ctx, arr, len(jobs), ctypes.byref(h)))
        self.h = h

    def launch(self, stream=0):
        self.q.check(self.q.lib.qtx_batch_launch(self.h, stream))

    def autotune(self, stream=0, repeats=2):
        t = BatchTuning()
        ms = ctypes.c_float()
        self.q.check(self.q.lib.qtx_batch_autotune(self.h, stream, repeats, ctypes.byref(t), ctypes.byref(ms)))
        return (t.workgroups_per_cu, t.unit_kops), ms.value

    @property
    def tuning(self):
        t = BatchTuning()
        self.q.check(self.q.lib.qtx_batch_get_tuning(self.h, ctypes.byref(t)))
        return (t.workgroups_per_cu, t.unit_kops)

    @tuning.setter
    def tuning(self, v):
        t = BatchTuning(*v)
        self.q.check(self.q.lib.qtx_batch_set_tuning(self.h, ctypes.byref(t)))

    def close(self):
        if self.h is not None:
            self.q.lib.qtx_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """A qtx_server: submit host-memory requests, wait for their results."""

    def __init__(self, q, device_bytes=0, staging_bytes=0, slots=0, max_batch_jobs=0, max_queue=0, unit_kops=0,
                 batch_bytes=0):
        self.q = q
        cfg = ServerConfig(device_bytes, staging_bytes, slots, max_batch_jobs, max_queue, unit_kops, batch_bytes)
        h = ctypes.c_void_p()
        q.check(q.lib.qtx_server_start(q.ctx, ctypes.byref(cfg), ctypes.byref(h)))
        self.h = h
        self._keep = {}

    def submit(self, j, src, dst):
        """Queue Job j reading host array src and writing host array dst; returns the ticket."""
        j.src = src.ctypes.data
        j.dst = dst.ctypes.data
        t = ctypes.c_uint64()
        self.q.check(self.q.lib.qtx_server_submit(self.h, ctypes.byref(j), dst.nbytes, ctypes.byref(t)))
        self._keep[t.value] = (src, dst)
        return t.value

    def wait(self, ticket, timeout_ms=-1):
        r = ServerResult()
        self.q.check(self.q.lib.qtx_server_wait(self.h, ticket, timeout_ms, ctypes.byref(r)))
        self._keep.pop(ticket, None)
        return r

    def cancel(self, ticket):
        """qtx_server_cancel: raises QtxError (status 5, EBUSY) if the request is already running."""
        self.q.check(self.q.lib.qtx_server_cancel(self.h, ticket))

    def stats(self):
        st = ServerStats()
        self.q.check(self.q.lib.qtx_server_stats_get(self.h, ctypes.byref(st)))
        return {n: getattr(st, n) for n, _ in st._fields_}

    def now(self):
        return self.q.lib.qtx_server_now(self.h)

    def stop(self):
        if self.h is not None:
            self.q.lib.qtx_server_stop(self.h)
            self.h = None
            self._keep.clear()


def mip_sizes(w, h, levels):
    """Sizes of the first `levels` levels of the chain (level 0 = (w, h))."""
    out = [(w, h)]
    for _ in range(levels - 1):
        w, h = max(w // 2, 1), max(h // 2, 1)
        out.append((w, h))
    return out


def mips_host(q, rgba, levels):
    """qtx_mips_host: levels 1.. of the chain of an (h, w, 4) uint8 array, as arrays."""
    rgba = np.ascontiguousarray(rgba, dtype=np.uint8)
    h, w = rgba.shape[:2]
    outs = [np.zeros((lh, lw, 4), np.uint8) for lw, lh in mip_sizes(w, h, levels)[1:]]
    n = max(len(outs), 1)
    ptrs = (ctypes.c_void_p * n)(*[o.ctypes.data for o in outs])
    caps = (ctypes.c_size_t * n)(*[o.nbytes for o in outs])
    q.check(q.lib.qtx_mips_host(q.ctx, _ptr(rgba), w, h, levels, ptrs, caps))
    return outs


def mips_encode_host(q, tmpl, rgba, levels):
    """qtx_mips_encode_host: the encoded levels 0.. of the chain, as uint8 arrays."""
    rgba = np.ascontiguousarray(rgba, dtype=np.uint8)
    h, w = rgba.shape[:2]
    bpb = 8 if tmpl.format in (FMT_BC1, FMT_BC4) else 16
    outs = [np.zeros(_nb(lw, lh) * bpb, np.uint8) for lw, lh in mip_sizes(w, h, levels)]
    ptrs = (ctypes.c_void_p * levels)(*[o.ctypes.data for o in outs])
    caps = (ctypes.c_size_t * levels)(*[o.nbytes for o in outs])
    q.check(q.lib.qtx_mips_encode_host(q.ctx, ctypes.byref(tmpl), _ptr(rgba), w, h, levels, ptrs, caps))
    return outs


def has_alpha_host(q, rgba):
    """qtx_has_alpha_host on an (h, w, 4) uint8 array."""
    rgba = np.ascontiguousarray(rgba, dtype=np.uint8)
    h, w = rgba.shape[:2]
    out = ctypes.c_int32(-1)
    q.check(q.lib.qtx_has_alpha_host(q.ctx, _ptr(rgba), w, h, ctypes.byref(out)))
    return bool(out.value)


def swizzle_bc3nm_host(q, rgba):
    """qtx_swizzle_bc3nm_host: (R, G, B, A) -> (255, G, G, R) of an (h, w, 4) uint8 array."""
    rgba = np.ascontiguousarray(rgba, dtype=np.uint8)
    h, w = rgba.shape[:2]
    out = np.empty_like(rgba)
    q.check(q.lib.qtx_swizzle_bc3nm_host(q.ctx, _ptr(rgba), w, h, _ptr(out)))
    return out


def mips_encode_auto_host(q, params, rgba, levels):
    """qtx_mips_encode_auto_host: (format, encoded levels) of the chain."""
    rgba = np.ascontiguousarray(rgba, dtype=np.uint8)
    h, w = rgba.shape[:2]
    outs = [np.zeros(_nb(lw, lh) * 16, np.uint8) for lw, lh in mip_sizes(w, h, levels)]
    ptrs = (ctypes.c_void_p * levels)(*[o.ctypes.data for o in outs])
    caps = (ctypes.c_size_t * levels)(*[o.nbytes for o in outs])
    fmt = ctypes.c_int32(0)
    q.check(q.lib.qtx_mips_encode_auto_host(q.ctx, ctypes.byref(params), _ptr(rgba), w, h, levels, ptrs, caps,
                                            ctypes.byref(fmt)))
    bpb = 16 if fmt.value == FMT_BC3 else 8
    return fmt.value, [o[:_nb(lw, lh) * bpb] for o, (lw, lh) in zip(outs, mip_sizes(w, h, levels))]


def run_host_batch(q, jobs):
    """qtx_batch_run_host over numpy buffers: jobs is a list of (Job, src_array, dst_array)."""
    arr = (Job * len(jobs))()
    caps = (ctypes.c_size_t * len(jobs))()
    for i, (j, src, dst) in enumerate(jobs):
        j.src = src.ctypes.data
        j.dst = dst.ctypes.data
        arr[i] = j
        caps[i] = dst.nbytes
    q.check(q.lib.qtx_batch_run_host(q.ctx, arr, caps, len(jobs)))


def declared_symbols(header=os.path.join(os.path.dirname(HERE), 'include', 'qtx.h')):
    """Every function name declared in include/qtx.h."""
    import re
    src = open(header).read()
    return sorted(set(re.findall(r'\b(qtx_[a-z0-9_]+)\s*\(', src)))
