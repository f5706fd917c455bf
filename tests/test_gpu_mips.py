"""GPU: mip chains (qtx_mip*), bit-exact against the oracle's restatement of
Pillow's BILINEAR resize (itself pinned against Pillow in test_oracle.py), and
the chain encode against per-level encodes through the plain entry points.
"""
import ctypes

import numpy as np
import pytest

from quicktex_amd import synth
from quicktex_amd.capi import (FMT_BC1, FMT_BC3, FMT_BC4, FMT_BC5, OP_ENCODE, QtxError, job, mip_sizes,
                               mips_encode_host, mips_host)

pytestmark = pytest.mark.gpu

# (96, 48), (256, 256), (2048, 16) and (4096, 8) take the horizontal pass's box sums
# at shrinks of 16 and more (96: 16 and 32, then a level that is no power-of-two
# shrink); (8192, 2) is past their width limit.
SIZES = [(1, 1), (1, 2), (2, 1), (3, 3), (5, 7), (13, 1), (1, 29), (37, 23), (64, 64), (96, 48), (255, 3),
         (256, 256), (1000, 600), (513, 257), (2048, 16), (4096, 8), (3, 2048), (9000, 3), (8192, 2)]


@pytest.mark.parametrize('w,h', SIZES)
def test_mips_host_match_oracle(qtx, oracle, w, h):
    rng = np.random.default_rng(w * 31 + h)
    img = rng.integers(0, 256, (h, w, 4), dtype=np.uint8)
    levels = qtx.lib.qtx_mip_count(w, h, 0)
    got = mips_host(qtx, img, levels)
    assert len(got) == levels - 1
    for m, (lw, lh) in enumerate(mip_sizes(w, h, levels)[1:], 1):
        want = oracle.resize(img, lw, lh)
        bad = np.argwhere(got[m - 1] != want)
        assert len(bad) == 0, (w, h, m, bad[:3].tolist())


def test_mip_count(qtx):
    L = qtx.lib
    assert L.qtx_mip_count(1, 1, 0) == 1
    assert L.qtx_mip_count(4096, 4096, 0) == 13
    assert L.qtx_mip_count(1000, 600, 0) == 10  # 1000 -> 1: floor(log2(1000)) + 1
    assert L.qtx_mip_count(1000, 600, 3) == 3
    assert L.qtx_mip_count(5, 1, 0) == 3
    assert L.qtx_mip_count(0, 5, 0) == 0


def test_mips_device_and_workspace(qtx, oracle):
    from quicktex_amd import hip
    w, h = 300, 170
    img = synth.mixture_rgba(w, h, seed=3)
    levels = qtx.lib.qtx_mip_count(w, h, 0)
    ws_bytes = qtx.lib.qtx_mips_workspace_bytes(w, h, levels)
    assert ws_bytes > 0
    d_in = hip.Buffer.from_array(img)
    sizes = mip_sizes(w, h, levels)[1:]
    outs = [hip.Buffer(lw * lh * 4) for lw, lh in sizes]
    ws = hip.Buffer(ws_bytes)
    ptrs = (ctypes.c_void_p * len(outs))(*[o.ptr for o in outs])
    s = hip.Stream()
    # a workspace one byte short is refused
    with pytest.raises(QtxError):
        qtx.check(qtx.lib.qtx_mips_device(qtx.ctx, d_in.ptr, w, h, levels, ptrs, ws.ptr, ws_bytes - 1, s.handle))
    qtx.check(qtx.lib.qtx_mips_device(qtx.ctx, d_in.ptr, w, h, levels, ptrs, ws.ptr, ws_bytes, s.handle))
    s.synchronize()
    for (lw, lh), o in zip(sizes, outs):
        got = o.download().reshape(lh, lw, 4)
        assert (got == oracle.resize(img, lw, lh)).all(), (lw, lh)
    s.destroy()


def test_mips_errors(qtx):
    img = np.zeros((8, 8, 4), np.uint8)
    with pytest.raises(QtxError):
        mips_host(qtx, img, 5)  # an 8x8 chain has 4 levels
    with pytest.raises(QtxError):
        mips_host(qtx, img, 0)
    assert qtx.lib.qtx_mips_workspace_bytes(8, 8, 9) == 0


def _tmpl(qtx, fmt, level=5, interp=0):
    return job(OP_ENCODE, fmt, 0, 0, 0, 0, params=qtx.params(level, 4, interp), chan0=0, chan1=1)


def _encode_level(qtx, fmt, img, p):
    if fmt == FMT_BC1:
        return qtx.bc1_encode(img, p)
    if fmt == FMT_BC3:
        return qtx.bc3_encode(img, p)
    if fmt == FMT_BC4:
        return qtx.bc4_encode(img, 0)
    return qtx.bc5_encode(img, 0, 1)


@pytest.mark.parametrize('fmt', [FMT_BC1, FMT_BC3, FMT_BC4, FMT_BC5])
@pytest.mark.parametrize('w,h', [(256, 256), (1000, 600), (37, 23), (1, 1)])
def test_mips_encode_matches_per_level(qtx, oracle, fmt, w, h):
    img = synth.mixture_rgba(w, h, seed=w + h)
    img[..., 3] = synth.alpha_channel(w, h, seed=w * h + 1)
    t = _tmpl(qtx, fmt, level=10)
    levels = qtx.lib.qtx_mip_count(w, h, 0)
    got = mips_encode_host(qtx, t, img, levels)
    assert len(got) == levels
    for m, (lw, lh) in enumerate(mip_sizes(w, h, levels)):
        level_img = img if m == 0 else oracle.resize(img, lw, lh)
        want = _encode_level(qtx, fmt, level_img, t.bc1)
        assert (got[m] == want).all(), (fmt, w, h, m)


def test_mips_encode_partial_chain(qtx, oracle):
    w, h = 640, 480
    img = synth.mixture_rgba(w, h, seed=9)
    t = _tmpl(qtx, FMT_BC1, level=18, interp=2)
    got = mips_encode_host(qtx, t, img, 3)
    assert len(got) == 3
    for m, (lw, lh) in enumerate(mip_sizes(w, h, 3)):
        level_img = img if m == 0 else oracle.resize(img, lw, lh)
        assert (got[m] == qtx.bc1_encode(level_img, t.bc1)).all(), m
