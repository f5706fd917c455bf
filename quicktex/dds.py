"""quicktex.dds (quicktex/dds.py), from quicktex_amd.dds."""
from quicktex_amd.dds import *  # noqa: F401,F403
from quicktex_amd.dds import (Caps0, DDSFile, DDSFlags, DDSFormat, PFFlags, dds_formats, encode,  # noqa: F401
                              encode_auto, read, swizzle_bc3nm)
