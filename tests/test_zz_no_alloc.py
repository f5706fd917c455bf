"""GPU: qtx_bc1_encode_device_noalloc, the BC1 encode that allocates nothing.

Called through the C ABI and compared, byte for byte, with qtx_bc1_encode_device
on the same input: two calls in a row on one stream, the second with a smaller
texture (the static scratch keeps the first call's contents), both at sizes and
settings that take the queued lane kernel and park endpoint searches (level 18),
then a setting that takes the static lane kernel over its block lists (kept in
the same scratch) and one below the lane crossovers (the 16-lane encoder; no
scratch is touched).  Sizes past the maximum are refused before any device
work.
"""
import numpy as np
import pytest

from quicktex_amd import synth
from quicktex_amd.capi import QtxError

pytestmark = pytest.mark.gpu

MAX_BLOCKS = 262144  # QTX_NOALLOC_MAX_BLOCKS


def _both(q, hip, img, p, stream):
    h, w = img.shape[:2]
    nb = ((w + 3) // 4) * ((h + 3) // 4)
    d_in = hip.Buffer.from_array(img)
    d_ref, d_got = hip.Buffer(nb * 8), hip.Buffer(nb * 8)
    d_got.zero(0xA5)
    q.bc1_encode_device(p, d_in.ptr, w, h, d_ref.ptr, stream.handle)
    q.bc1_encode_device_noalloc(p, d_in.ptr, w, h, d_got.ptr, stream.handle)
    stream.synchronize()
    return d_ref.download().reshape(-1, 8), d_got.download().reshape(-1, 8)


def test_noalloc_equals_device_entry_point(qtx):
    from quicktex_amd import hip
    s = hip.Stream()
    try:
        cases = [((2048, 2048), 18, 0x13, 2),  # the maximum: queued kernel, parked searches
                 ((1798, 1602), 18, 0x13, 0),  # smaller, ragged: the scratch holds the last call's words
                 ((1024, 512), 9, 4, 1),       # static lane kernel over block lists
                 ((256, 64), 18, 0x13, 3)]     # 16-lane encoder
        for (w, h), level, mode, interp in cases:
            print('noalloc case %dx%d level %d mode %#x interp %d' % (w, h, level, mode, interp), flush=True)
            img = synth.mixture_rgba(w, h, seed=0x0A110C + w + h, near_black=0.1)
            p = qtx.params(level, mode, interp)
            ref, got = _both(qtx, hip, img, p, s)
            bad = np.nonzero((ref != got).any(axis=1))[0]
            assert len(bad) == 0, '%dx%d: %d blocks differ, first %s' % (w, h, len(bad), bad[:4].tolist())
    finally:
        s.destroy()


def test_noalloc_refuses_past_maximum(qtx):
    from quicktex_amd import hip
    p = qtx.params(18, 0x13, 0)
    d_in, d_out = hip.Buffer(64), hip.Buffer(64)  # never touched: the size is refused on the host first
    with pytest.raises(QtxError) as e:
        qtx.bc1_encode_device_noalloc(p, d_in.ptr, 2052, 2048, d_out.ptr, 0)  # 513 x 512 blocks
    assert e.value.status == 2 and 'at most %d' % MAX_BLOCKS in str(e.value)
