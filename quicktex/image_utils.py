"""quicktex.image_utils (quicktex/image_utils.py), from quicktex_amd.image_utils."""
from quicktex_amd.image_utils import mip_sizes, resize_no_premultiply  # noqa: F401
