"""GPU: the host entry points' staging (quicktex_amd/csrc/staging.cpp).

Host buffers of 4 MiB and more move through the context's two pinned 8 MiB
chunks, a few threads copying one chunk while the DMA engine moves the other;
smaller ones take the runtime's own path.  Checked here, byte for byte, against
the device entry points on the same data: inputs just below and at the threshold,
of two chunks with a short last one and of several chunks at ragged sizes, for
encodes (the staged upload) and decodes (the staged download), and calls of
different sizes in a row (a chunk's next use waits for its last DMA).
"""
import numpy as np
import pytest

from quicktex_amd import synth

pytestmark = pytest.mark.gpu

SIZES = [(1024, 1023), (1024, 1024), (2048, 1025), (2501, 3001)]  # 4 MiB - 4 KiB, 4 MiB, 2 chunks, 4 chunks


def _device_bc4(q, hip, img, channel):
    h, w = img.shape[:2]
    nb = ((w + 3) // 4) * ((h + 3) // 4)
    d_in, d_out = hip.Buffer.from_array(img), hip.Buffer(nb * 8)
    q.bc4_encode_device(channel, d_in.ptr, w, h, d_out.ptr)
    hip.synchronize()
    return d_out.download().reshape(-1)


def _device_bc1_decode(q, hip, blocks, w, h):
    d_in, d_out = hip.Buffer.from_array(blocks), hip.Buffer(w * h * 4)
    q.bc1_decode_device(0, 1, d_in.ptr, w, h, d_out.ptr)
    hip.synchronize()
    return d_out.download().reshape(h, w, 4)


@pytest.mark.parametrize('w,h', SIZES)
def test_staged_encode(qtx, w, h):
    from quicktex_amd import hip
    img = synth.mixture_rgba(w, h, seed=900 + w, near_black=0.1)
    assert (qtx.bc4_encode(img, 1) == _device_bc4(qtx, hip, img, 1)).all()
    p = qtx.params(0, 4, 0)
    d_in = hip.Buffer.from_array(img)
    nb = ((w + 3) // 4) * ((h + 3) // 4)
    d_out = hip.Buffer(nb * 8)
    qtx.bc1_encode_device(p, d_in.ptr, w, h, d_out.ptr)
    hip.synchronize()
    assert (qtx.bc1_encode(img, p) == d_out.download().reshape(-1)).all()


@pytest.mark.parametrize('w,h', SIZES)
def test_staged_decode(qtx, w, h):
    from quicktex_amd import hip
    rng = np.random.default_rng(w * h)
    nb = ((w + 3) // 4) * ((h + 3) // 4)
    blocks = rng.integers(0, 256, nb * 8, dtype=np.uint8)
    got = qtx.bc1_decode(blocks, w, h, 0, True)
    assert (got == _device_bc1_decode(qtx, hip, blocks, w, h)).all()


def test_calls_in_a_row(qtx):
    """Large, small and large again, encodes and decodes interleaved, repeated."""
    from quicktex_amd import hip
    big = synth.mixture_rgba(3000, 2000, seed=911, near_black=0.1)
    small = synth.mixture_rgba(300, 200, seed=912, near_black=0.1)
    want_big, want_small = _device_bc4(qtx, hip, big, 0), _device_bc4(qtx, hip, small, 0)
    rng = np.random.default_rng(913)
    blocks = rng.integers(0, 256, 750 * 500 * 8, dtype=np.uint8)
    want_dec = _device_bc1_decode(qtx, hip, blocks, 3000, 2000)
    for _ in range(3):
        assert (qtx.bc4_encode(big, 0) == want_big).all()
        assert (qtx.bc1_decode(blocks, 3000, 2000, 0, True) == want_dec).all()
        assert (qtx.bc4_encode(small, 0) == want_small).all()
