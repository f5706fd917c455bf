"""Mip-chain helpers with the names and contract of quicktex.image_utils.

``mip_sizes`` gives the level sizes the reference's DDS writer and reader use
(quicktex/image_utils.py:9-35): level k of a w x h texture is
max(w >> k, 1) x max(h >> k, 1), and the full chain ends with the first 1 x 1
level, which is level bit_length(max(w, h)) - 1.

``resize_no_premultiply`` is the single-image Pillow resize of
quicktex/image_utils.py:38-52 (colour and alpha resized as separate bands, so
that Pillow does not premultiply).  The DDS encoder here does not call it:
:func:`quicktex_amd.dds.encode` resamples every level on the GPU with the same
arithmetic (include/qtx.h, qtx_mips_*).
"""
from typing import List, Optional, Tuple


def chain_length(dimensions: Tuple[int, ...]) -> int:
    """Levels from ``dimensions`` down to 1 x 1 inclusive."""
    return max(dimensions).bit_length()


def mip_sizes(dimensions: Tuple[int, int], mip_count: Optional[int] = None) -> List[Tuple[int, int]]:
    """Sizes of the mip chain of a texture, level 0 (``dimensions``) first.

    Every level halves the sides of the one before, rounding down and stopping
    at 1.  ``mip_count`` (None or 0: the whole chain) caps the number of levels;
    a chain reaches 1 x 1 at most once.  Non-positive sides or a negative
    ``mip_count`` raise AssertionError, as the reference's checks do.
    """
    if min(dimensions) <= 0:
        raise AssertionError('Invalid source dimensions %r: every side must be positive' % (tuple(dimensions),))
    levels = chain_length(dimensions)
    if mip_count:
        if mip_count < 0:
            raise AssertionError('mip_count must be greater than 0, not %d' % mip_count)
        levels = min(levels, mip_count)
    return [tuple(max(side >> k, 1) for side in dimensions) for k in range(levels)]


def resize_no_premultiply(image, size: Tuple[int, int]):
    """Pillow BILINEAR resize to ``size``; an RGBA image has its colour bands and its
    alpha band resized apart and joined again, so alpha does not weight colour."""
    from PIL import Image
    if image.mode != 'RGBA':
        return image.resize(size, Image.BILINEAR)
    colour = image.convert('RGB').resize(size, Image.BILINEAR)
    colour.putalpha(image.getchannel('A').resize(size, Image.BILINEAR))
    return colour
