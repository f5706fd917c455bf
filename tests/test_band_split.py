"""CPU: the row-band decomposition of the large host encodes, on the oracle.

The BC1/BC3 host entry points encode a large texture as bands of whole block
rows plus, for a height that is not a multiple of 4, its last block row on its
own, as a texture of height 4 whose rows are the ones the reference's padding
reads: texel row y of the texture is row y % h (quicktex/Texture.h:107, the
block fetch's wrap-around), so the missing rows come from the top of the
texture.  Here the oracle encodes each unit separately and the concatenation
must equal its encode of the whole texture; a unit that clamped instead of
wrapping, or a band cut inside a block row, differs in the last block row.
"""
import numpy as np
import pytest

from quicktex_amd import synth
from tests._cpu_libs import oracle


def _units(img, plan):
    h = img.shape[0]
    full = h // 4
    out, r0, acc = [], 0, 0
    for part in plan:
        acc += part
        r1 = full * acc // 16
        if r1 > r0:
            out.append(np.ascontiguousarray(img[4 * r0:4 * r1]))
        r0 = max(r0, r1)
    if 4 * full < h:
        out.append(np.ascontiguousarray(np.stack([img[y if y < h else y - h] for y in range(4 * full, 4 * full + 4)])))
    return out


@pytest.mark.parametrize('w,h', [(67, 130), (33, 41), (130, 66), (45, 64)])
@pytest.mark.parametrize('plan', [(2, 6, 8), (4, 12), (1, 3, 4, 8)])
def test_bands_equal_whole(w, h, plan):
    o = oracle()
    img = synth.mixture_rgba(w, h, seed=0xBA + w * h, near_black=0.1)
    for level, bc3 in ((0, False), (5, True)):
        p = o.params(level, 4, 0)
        enc = o.bc3_encode if bc3 else o.bc1_encode
        whole = enc(img, p).reshape(-1)
        parts = np.concatenate([enc(u, p).reshape(-1) for u in _units(img, plan)])
        assert np.array_equal(whole, parts), (w, h, plan, level)


def test_clamped_padding_differs():
    """The check has teeth: padding the last block row by clamping to the last
    texel row (instead of wrapping) changes its blocks."""
    o = oracle()
    w, h = 67, 130
    img = synth.mixture_rgba(w, h, seed=0xBB, near_black=0.1)
    p = o.params(5, 4, 0)
    whole = o.bc1_encode(img, p).reshape(-1)
    full = h // 4
    clamped = np.stack([img[min(y, h - 1)] for y in range(4 * full, 4 * full + 4)])
    tail = o.bc1_encode(np.ascontiguousarray(clamped), p).reshape(-1)
    assert not np.array_equal(whole[-tail.size:], tail)
