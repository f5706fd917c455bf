"""Mip-chain helpers (quicktex/image_utils.py).

``mip_sizes`` is the reference's size rule.  ``resize_no_premultiply`` is the
reference's single-image resize through Pillow; the DDS encoder does not call
it: :func:`quicktex_amd.dds.encode` resamples every level of the chain on the
GPU with the same arithmetic (bit-exact, include/qtx.h qtx_mips_*).
"""
import math
from typing import List, Optional, Tuple


def mip_sizes(dimensions: Tuple[int, int], mip_count: Optional[int] = None) -> List[Tuple[int, int]]:
    """Sizes of a mip chain, level 0 first (quicktex/image_utils.py:9-38): each
    level halves both sides of the one before, rounding down, at least 1; the
    chain ends at 1x1 or after ``mip_count`` levels (default: the full chain)."""
    assert all([dim > 0 for dim in dimensions]), "Invalid source dimensions"
    if not mip_count:
        mip_count = math.ceil(math.log2(max(dimensions)) + 1)
    assert mip_count > 0, "mip_count must be greater than 0"
    chain = []
    for _ in range(mip_count):
        chain.append(dimensions)
        if all([dim == 1 for dim in dimensions]):
            break
        dimensions = tuple([max(dim // 2, 1) for dim in dimensions])
    return chain


def resize_no_premultiply(image, size: Tuple[int, int]):
    """Pillow BILINEAR resize of an image, RGBA without premultiplied alpha
    (quicktex/image_utils.py:42-52): the RGB bands and the alpha band separately."""
    from PIL import Image
    if image.mode == 'RGBA':
        rgb = image.convert('RGB').resize(size, Image.BILINEAR)
        rgb.putalpha(image.getchannel('A').resize(size, Image.BILINEAR))
        return rgb
    return image.resize(size, Image.BILINEAR)
