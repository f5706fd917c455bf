"""GPU: the host entry points' staging and row bands (quicktex_amd/csrc/staging.cpp,
quicktex_amd/csrc/capi.cpp).

Pageable host buffers move through the context's eight pinned 8 MiB chunks, a few
threads copying one chunk while the DMA engine moves the others: uploads of 8 MiB
and more in whole chunks, downloads of 2 MiB and more (below 32 MiB in four smaller
chunks); smaller ones take the runtime's own path, and pinned ones go to the DMA
engine directly.  BC1 and BC3 host encodes of 16 MiB of texels and more are
cut into bands of block rows, each encoded as soon as its rows have landed, on
two streams in turn.  Checked here, byte for byte, against the device entry
points on the same data: inputs just below and at the staging threshold, of
several chunks with a short last one, banded encodes at ragged sizes, of one
block row's height per band and of wide, short textures, at cheap and expensive
levels, from pageable and pinned memory; decodes (the staged download); an
error raised by a band's launch; calls of different sizes in a row (a chunk's
next use waits for its last DMA); and threads sharing a context beside threads
with contexts of their own.
"""
import ctypes

import numpy as np
import pytest

from quicktex_amd import synth
from quicktex_amd.capi import QtxError, _ptr

pytestmark = pytest.mark.gpu

# uploads: 8 MiB - 4 KiB, 8 MiB (one chunk), two chunks with a short last one, four ragged
# chunks, nine (more chunks than slots: the first slot comes round again within the call);
# downloads (the decodes' texels): 2 MiB - 2 KiB and 2 MiB (the threshold), and the same
# sizes in quarter chunks (below 32 MiB) or whole ones
SIZES = [(512, 1023), (512, 1024), (1024, 2047), (1024, 2048), (2048, 1025), (2501, 3001), (4200, 4100)]
# banded BC1/BC3 encodes: 16 MiB exactly, ragged both ways, wide and short (65
# block rows), tall and narrow
BANDED = [(2048, 2048), (2050, 2047), (16384, 257), (1024, 4100)]


def _nb(w, h):
    return ((w + 3) // 4) * ((h + 3) // 4)


def _device_bc4(q, hip, img, channel):
    h, w = img.shape[:2]
    d_in, d_out = hip.Buffer.from_array(img), hip.Buffer(_nb(w, h) * 8)
    q.bc4_encode_device(channel, d_in.ptr, w, h, d_out.ptr)
    hip.synchronize()
    return d_out.download().reshape(-1)


def _device_colour(q, hip, img, p, bc3):
    h, w = img.shape[:2]
    d_in, d_out = hip.Buffer.from_array(img), hip.Buffer(_nb(w, h) * (16 if bc3 else 8))
    (q.bc3_encode_device if bc3 else q.bc1_encode_device)(p, d_in.ptr, w, h, d_out.ptr)
    hip.synchronize()
    return d_out.download().reshape(-1)


def _device_bc1_decode(q, hip, blocks, w, h):
    d_in, d_out = hip.Buffer.from_array(blocks), hip.Buffer(w * h * 4)
    q.bc1_decode_device(0, 1, d_in.ptr, w, h, d_out.ptr)
    hip.synchronize()
    return d_out.download().reshape(h, w, 4)


def _first_bad(got, want, bs):
    bad = np.nonzero((got.reshape(-1, bs) != want.reshape(-1, bs)).any(axis=1))[0]
    return len(bad), bad[:4].tolist()


@pytest.mark.parametrize('w,h', SIZES)
def test_staged_encode(qtx, w, h):
    from quicktex_amd import hip
    img = synth.mixture_rgba(w, h, seed=900 + w, near_black=0.1)
    assert (qtx.bc4_encode(img, 1) == _device_bc4(qtx, hip, img, 1)).all()
    p = qtx.params(0, 4, 0)
    assert (qtx.bc1_encode(img, p) == _device_colour(qtx, hip, img, p, False)).all()


@pytest.mark.parametrize('w,h', SIZES)
def test_staged_decode(qtx, w, h):
    from quicktex_amd import hip
    rng = np.random.default_rng(w * h)
    blocks = rng.integers(0, 256, _nb(w, h) * 8, dtype=np.uint8)
    got = qtx.bc1_decode(blocks, w, h, 0, True)
    assert (got == _device_bc1_decode(qtx, hip, blocks, w, h)).all()


@pytest.mark.parametrize('w,h', BANDED)
@pytest.mark.parametrize('level,mode,interp,bc3', [(18, 0x13, 0, False), (5, 4, 2, False), (10, 4, 3, True)])
def test_banded_encode(qtx, w, h, level, mode, interp, bc3):
    from quicktex_amd import hip
    print('banded %dx%d level %d bc3 %d' % (w, h, level, bc3), flush=True)
    img = synth.mixture_rgba(w, h, seed=0xBA4D + w + h, near_black=0.1)
    p = qtx.params(level, mode, interp)
    want = _device_colour(qtx, hip, img, p, bc3)
    got = (qtx.bc3_encode if bc3 else qtx.bc1_encode)(img, p)
    n, first = _first_bad(got, want, 16 if bc3 else 8)
    assert n == 0, '%dx%d: %d blocks differ, first %s' % (w, h, n, first)


def test_banded_encode_pinned(qtx):
    """Pinned input and output: the bands' uploads and downloads go to the DMA engine directly."""
    from quicktex_amd import hip
    w, h = 2500, 2100
    img = synth.mixture_rgba(w, h, seed=0xBA4E, near_black=0.1)
    p = qtx.params(12, 4, 1)
    want = _device_colour(qtx, hip, img, p, True)
    src, dst = hip.Pinned(img.nbytes), hip.Pinned(want.nbytes)
    src.array[:] = img.reshape(-1)
    dst.array[:] = 0xA5
    qtx.check(qtx.lib.qtx_bc3_encode(qtx.ctx, ctypes.byref(p), _ptr(src.array), w, h, _ptr(dst.array), dst.nbytes))
    n, first = _first_bad(dst.array, want, 16)
    assert n == 0, '%d blocks differ, first %s' % (n, first)


def test_banded_encode_error_then_ok(qtx):
    """A launch refused in the first band fails the call; the context works on."""
    w, h = 2048, 2048
    img = synth.mixture_rgba(w, h, seed=0xBA4F, near_black=0.1)
    p = qtx.params(5, 4, 0)
    p.interpolator = 9
    with pytest.raises(QtxError):
        qtx.bc1_encode(img, p)
    from quicktex_amd import hip
    p = qtx.params(5, 4, 0)
    assert (qtx.bc1_encode(img, p) == _device_colour(qtx, hip, img, p, False)).all()


def test_calls_in_a_row(qtx):
    """Large, small and large again, banded and plain encodes and decodes interleaved, repeated."""
    from quicktex_amd import hip
    big = synth.mixture_rgba(3000, 2000, seed=911, near_black=0.1)
    small = synth.mixture_rgba(300, 200, seed=912, near_black=0.1)
    want_big, want_small = _device_bc4(qtx, hip, big, 0), _device_bc4(qtx, hip, small, 0)
    p = qtx.params(8, 4, 0)
    want_colour = _device_colour(qtx, hip, big, p, False)
    rng = np.random.default_rng(913)
    blocks = rng.integers(0, 256, 750 * 500 * 8, dtype=np.uint8)
    want_dec = _device_bc1_decode(qtx, hip, blocks, 3000, 2000)
    for _ in range(3):
        assert (qtx.bc4_encode(big, 0) == want_big).all()
        assert (qtx.bc1_encode(big, p) == want_colour).all()
        assert (qtx.bc1_decode(blocks, 3000, 2000, 0, True) == want_dec).all()
        assert (qtx.bc4_encode(small, 0) == want_small).all()


def test_threads_and_contexts(qtx):
    """Four threads share one context (its host lock serialises the staging and the
    bands) while two more each use a context of their own, all with large banded
    encodes, staged decodes and small calls; every result equals the device path's."""
    import threading

    from quicktex_amd import hip
    from tests._qtx import Qtx
    imgs = [synth.mixture_rgba(2100 + 60 * k, 2050, seed=0x7A + k, near_black=0.1) for k in range(3)]
    small = synth.mixture_rgba(200, 120, seed=0x7F, near_black=0.1)
    p = qtx.params(9, 4, 1)
    want = [_device_colour(qtx, hip, im, p, False) for im in imgs]
    want_small = _device_colour(qtx, hip, small, p, False)
    rng = np.random.default_rng(0x7B)
    blocks = rng.integers(0, 256, 600 * 520 * 8, dtype=np.uint8)
    want_dec = _device_bc1_decode(qtx, hip, blocks, 2400, 2080)
    errors = []

    def work(q, seed):
        try:
            order = np.random.default_rng(seed).permutation(6)
            for i in order:
                if i < 3:
                    assert (q.bc1_encode(imgs[i], p) == want[i]).all(), ('encode', i)
                elif i == 3:
                    assert (q.bc1_decode(blocks, 2400, 2080, 0, True) == want_dec).all(), 'decode'
                else:
                    assert (q.bc1_encode(small, p) == want_small).all(), 'small'
        except Exception as e:  # reported on the main thread
            errors.append(repr(e))

    own = [Qtx(), Qtx()]
    try:
        threads = [threading.Thread(target=work, args=(qtx, 100 + k)) for k in range(4)]
        threads += [threading.Thread(target=work, args=(q, 200 + k)) for k, q in enumerate(own)]
        for t in threads:
            t.start()
        for t in threads:
            t.join(120)
        assert not any(t.is_alive() for t in threads), 'a thread did not finish'
    finally:
        for q in own:
            q.close()
    assert not errors, errors


def _device_call(q, hip, fn, src, out_bytes):
    """Runs a device entry point fn(d_src, d_out) on copies of src; returns its output bytes."""
    d_in, d_out = hip.Buffer.from_array(np.ascontiguousarray(src)), hip.Buffer(out_bytes)
    fn(d_in.ptr, d_out.ptr)
    hip.synchronize()
    return d_out.download().reshape(-1)


# Host calls on pageable buffers whose input and output take at most 1.5 MiB together run
# their kernels on the context's pinned host memory (no DMA transfer): one block, ragged, small, and either side of the
# limit -- BC1 encode and decode 590x590 in place, 592x592 through the DMA engine; BC3/BC5
# encode 560x560 in place, 590x590 through the DMA engine
DIRECT = [(1, 1), (5, 3), (130, 66), (560, 560), (590, 590), (592, 592)]


@pytest.mark.parametrize('w,h', DIRECT)
def test_direct_calls(qtx, w, h):
    """Every host entry point, against its device entry point on the same data."""
    from quicktex_amd import hip
    print('direct %dx%d' % (w, h), flush=True)
    img = synth.mixture_rgba(w, h, seed=0xD1 + w * 7 + h, near_black=0.1)
    nb = _nb(w, h)
    p5, p10 = qtx.params(5, 4, 0), qtx.params(10, 4, 2)
    dev = lambda fn, src, n: _device_call(qtx, hip, fn, src, n)  # noqa: E731
    assert (qtx.bc1_encode(img, p5) == dev(lambda i, o: qtx.bc1_encode_device(p5, i, w, h, o), img, nb * 8)).all()
    assert (qtx.bc3_encode(img, p10) == dev(lambda i, o: qtx.bc3_encode_device(p10, i, w, h, o), img, nb * 16)).all()
    assert (qtx.bc4_encode(img, 1) == dev(lambda i, o: qtx.bc4_encode_device(1, i, w, h, o), img, nb * 8)).all()
    assert (qtx.bc5_encode(img, 0, 2) == dev(lambda i, o: qtx.bc5_encode_device(0, 2, i, w, h, o), img, nb * 16)).all()
    rng = np.random.default_rng(w * 1000 + h)
    b8, b16 = rng.integers(0, 256, nb * 8, dtype=np.uint8), rng.integers(0, 256, nb * 16, dtype=np.uint8)
    px = w * h * 4
    assert (qtx.bc1_decode(b8, w, h, 3, True).reshape(-1)
            == dev(lambda i, o: qtx.bc1_decode_device(3, 1, i, w, h, o), b8, px)).all()
    assert (qtx.bc3_decode(b16, w, h, 1).reshape(-1) == dev(lambda i, o: qtx.bc3_decode_device(1, i, w, h, o), b16, px)).all()
    assert (qtx.bc4_decode(b8, w, h, 2).reshape(-1) == dev(lambda i, o: qtx.bc4_decode_device(2, i, w, h, o), b8, px)).all()
    assert (qtx.bc5_decode(b16, w, h, 1, 3).reshape(-1)
            == dev(lambda i, o: qtx.bc5_decode_device(1, 3, i, w, h, o), b16, px)).all()


def test_direct_error_then_ok(qtx):
    """A launch refused on the in-place path fails the call; the context works on."""
    from quicktex_amd import hip
    img = synth.mixture_rgba(64, 48, seed=0xD2, near_black=0.1)
    p = qtx.params(5, 4, 0)
    p.interpolator = 9
    with pytest.raises(QtxError):
        qtx.bc1_encode(img, p)
    p = qtx.params(5, 4, 0)
    assert (qtx.bc1_encode(img, p) == _device_colour(qtx, hip, img, p, False)).all()


def test_direct_pinned(qtx):
    """Pinned input and output below the limit: the kernels read and write them in place."""
    from quicktex_amd import hip
    w, h = 200, 150
    img = synth.mixture_rgba(w, h, seed=0xD3, near_black=0.1)
    p = qtx.params(9, 4, 3)
    want = _device_colour(qtx, hip, img, p, True)
    src, dst = hip.Pinned(img.nbytes), hip.Pinned(want.nbytes)
    src.array[:] = img.reshape(-1)
    dst.array[:] = 0xA5
    qtx.check(qtx.lib.qtx_bc3_encode(qtx.ctx, ctypes.byref(p), _ptr(src.array), w, h, _ptr(dst.array), dst.nbytes))
    n, first = _first_bad(dst.array, want, 16)
    assert n == 0, '%d blocks differ, first %s' % (n, first)


@pytest.mark.parametrize('pin_src,pin_dst', [(True, True), (True, False), (False, True)])
def test_direct_pinned_decode(qtx, pin_src, pin_dst):
    """Small BC1 / BC5 decodes from and to pinned buffers (in place when both are), twice in a row."""
    from quicktex_amd import hip
    w, h = 300, 222
    nb = _nb(w, h)
    rng = np.random.default_rng(0xD4)
    for fmt, bpb in ((1, 8), (5, 16)):
        blocks = rng.integers(0, 256, nb * bpb, dtype=np.uint8)
        if fmt == 1:
            want = _device_call(qtx, hip, lambda i, o: qtx.bc1_decode_device(2, 1, i, w, h, o), blocks, w * h * 4)
        else:
            want = _device_call(qtx, hip, lambda i, o: qtx.bc5_decode_device(3, 0, i, w, h, o), blocks, w * h * 4)
        src = hip.Pinned(blocks.nbytes) if pin_src else None
        dst = hip.Pinned(w * h * 4) if pin_dst else None
        s_arr = src.array if src else blocks.copy()
        d_arr = dst.array if dst else np.zeros(w * h * 4, np.uint8)
        if src:
            s_arr[:] = blocks
        for _ in range(2):
            d_arr[:] = 0x5A
            if fmt == 1:
                qtx.check(qtx.lib.qtx_bc1_decode(qtx.ctx, 2, 1, _ptr(s_arr), w, h, _ptr(d_arr), d_arr.size))
            else:
                qtx.check(qtx.lib.qtx_bc5_decode(qtx.ctx, 3, 0, _ptr(s_arr), w, h, _ptr(d_arr), d_arr.size))
            assert (d_arr == want).all()
