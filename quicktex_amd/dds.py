"""DDS files of S3TC textures (quicktex/dds.py), with the mip chain built and
encoded on the GPU.

The container format is Microsoft's DDS (a 4-byte magic, a 124-byte header, then
each mip level's blocks); the header fields, flags and the FourCC codes of the
four formats are the reference's (dds.py:15-123).  Differences are in how the
work runs, not in the bytes:

* :func:`encode` converts the image with Pillow as the reference does
  (dds.py:276-279), then resamples every mip level from the base and encodes
  all of them in one batched GPU launch (include/qtx.h, qtx_mips_encode_host);
  the levels equal Pillow's BILINEAR resize (image_utils.resize_no_premultiply)
  and the blocks equal per-level encoder.encode calls.
* :meth:`DDSFile.decode_all` decodes every level in one batched launch.
* :func:`encode_auto` and :func:`swizzle_bc3nm` are the per-image choices of the
  reference's ``quicktex encode auto`` / ``encode bc3nm`` commands
  (cli/encode.py:52-55, 103-141), with the alpha test and the swizzle on the GPU.
"""
from __future__ import annotations

import enum
import os
import struct
import typing

from . import _native
from .image_utils import mip_sizes
from .s3tc import bc1, bc3, bc4, bc5


class DDSFormat:
    def __init__(self, name: str, texture, encoder, decoder, four_cc: str = None):
        self.four_cc = four_cc
        self.decoder = decoder
        self.encoder = encoder
        self.texture = texture
        self.name = name


dds_formats = [
    DDSFormat('BC1', bc1.BC1Texture, bc1.BC1Encoder, bc1.BC1Decoder, 'DXT1'),
    DDSFormat('BC3', bc3.BC3Texture, bc3.BC3Encoder, bc3.BC3Decoder, 'DXT5'),
    DDSFormat('BC4', bc4.BC4Texture, bc4.BC4Encoder, bc4.BC4Decoder, 'ATI1'),
    DDSFormat('BC5', bc5.BC5Texture, bc5.BC5Encoder, bc5.BC5Decoder, 'ATI2'),
]


class PFFlags(enum.IntFlag):
    """Pixel format flags (DDS_PIXELFORMAT.dwFlags)."""
    ALPHAPIXELS = 0x1
    ALPHA = 0x2
    FOURCC = 0x4
    RGB = 0x40
    YUV = 0x200
    LUMINANCE = 0x20000


class DDSFlags(enum.IntFlag):
    """Header flags (DDS_HEADER.dwFlags): which members hold valid data."""
    CAPS = 0x1
    HEIGHT = 0x2
    WIDTH = 0x4
    PITCH = 0x8
    PIXEL_FORMAT = 0x1000
    MIPMAPCOUNT = 0x20000
    LINEAR_SIZE = 0x80000
    DEPTH = 0x800000
    TEXTURE = CAPS | HEIGHT | WIDTH | PIXEL_FORMAT


class Caps0(enum.IntFlag):
    """Surface complexity (DDS_HEADER.dwCaps)."""
    COMPLEX = 0x8
    MIPMAP = 0x400000
    TEXTURE = 0x1000


@typing.final
class DDSFile:
    """A DDS file: header fields and one block texture per mip level."""

    magic = b'DDS '
    extension = 'dds'
    header_bytes = 124

    def __init__(self):
        self.flags: DDSFlags = DDSFlags.TEXTURE
        self.size: typing.Tuple[int, int] = (0, 0)
        self.pitch: int = 0
        self.depth: int = 1
        self.mipmap_count: int = 1
        self.pf_flags: PFFlags = PFFlags.FOURCC
        self.four_cc: str = "NONE"
        self.pixel_size: int = 0
        self.pixel_bitmasks: typing.Tuple[int, int, int, int] = (0, 0, 0, 0)
        self.caps: typing.Tuple[Caps0, int, int, int] = (Caps0.TEXTURE, 0, 0, 0)
        self.textures: typing.List = []
        self.format: DDSFormat = DDSFormat('NONE', None, None, None)

    def save(self, path: os.PathLike) -> None:
        """Writes the file: magic, header (dds.py:162-196 layout), every level's blocks."""
        header = struct.pack('<7I44x', DDSFile.header_bytes, int(self.flags), self.size[1], self.size[0], self.pitch,
                             self.depth, self.mipmap_count)
        header += struct.pack('<2I4s5I', 32, int(self.pf_flags), bytes(self.four_cc, 'ascii'), self.pixel_size,
                              *self.pixel_bitmasks)
        header += struct.pack('<4I4x', *[int(c) for c in self.caps])
        assert len(header) == DDSFile.header_bytes, 'error writing file: incorrect header size'
        with open(path, 'wb') as file:
            file.write(DDSFile.magic)
            file.write(header)
            for texture in self.textures:
                file.write(texture)

    def decode(self, mip: int = 0, *args, **kwargs):
        """One level as a Pillow RGBA image (decoder arguments as the format's decoder takes)."""
        from PIL import Image
        decoder = self.format.decoder(*args, **kwargs)
        texture = decoder.decode(self.textures[mip])
        return Image.frombuffer('RGBA', texture.size, texture)

    def decode_all(self, *args, **kwargs):
        """Every level as Pillow RGBA images, decoded in one batched launch."""
        from PIL import Image
        decoder = self.format.decoder(*args, **kwargs)
        textures = _native()._decode_all(decoder, list(self.textures))
        return [Image.frombuffer('RGBA', tex.size, tex) for tex in textures]


def read(path: os.PathLike) -> DDSFile:
    """Parses a DDS file of one of the four formats (dds.py:219-272 behaviour)."""
    with open(path, 'rb') as file:
        assert file.read(4) == DDSFile.magic, "Incorrect magic bytes in DDS file."
        dds = DDSFile()
        header_bytes = struct.unpack('<I', file.read(4))[0]
        assert header_bytes == DDSFile.header_bytes, "Incorrect DDS header size."
        dds.flags = DDSFlags(struct.unpack('<I', file.read(4))[0])
        dds.size = struct.unpack('<2I', file.read(8))[::-1]
        dds.pitch, dds.depth, dds.mipmap_count = struct.unpack('<3I', file.read(12))
        file.read(44)
        assert struct.unpack('<I', file.read(4))[0] == 32, "Incorrect pixel format size."
        dds.pf_flags = PFFlags(struct.unpack('<I', file.read(4))[0])
        dds.four_cc = file.read(4).decode()
        dds.pixel_size, *pixel_bitmasks = struct.unpack('<5I', file.read(20))
        dds.pixel_bitmasks = tuple(pixel_bitmasks)
        dds.caps = struct.unpack('<4I', file.read(16))
        file.read(4)
        assert file.tell() == 4 + DDSFile.header_bytes, "Unexpected EOF"

        if DDSFlags.DEPTH not in dds.flags:
            dds.depth = 1
        if DDSFlags.MIPMAPCOUNT not in dds.flags:
            dds.mipmap_count = 1
        if PFFlags.FOURCC not in dds.pf_flags:
            dds.four_cc = 'NONE'
        if dds.four_cc == 'DX10':
            raise NotImplementedError('DX10 headers are not yet supported')

        dds.format = next(entry for entry in dds_formats if entry.four_cc == dds.four_cc)
        dds.textures = []
        for size in mip_sizes(dds.size, dds.mipmap_count):
            texture = dds.format.texture(*size)
            nbytes = file.readinto(texture)
            assert nbytes == texture.nbytes, 'Unexpected end of file'
            dds.textures.append(texture)
        return dds


def encode(image, encoder, four_cc: str, mip_count: typing.Optional[int] = None) -> DDSFile:
    """A DDSFile of the image and its mip chain (dds.py:275-303 behaviour): the
    image is converted to RGBA (or RGBX without alpha) by Pillow, then every
    level is resampled and encoded on the GPU in one batch."""
    mode = 'RGBA' if 'A' in image.mode else 'RGBX'
    image.apply_transparency()
    image = image.convert(mode)

    sizes = mip_sizes(image.size, mip_count)
    data = image.tobytes('raw', mode)
    dds = DDSFile()
    dds.textures = list(_native()._encode_mips(encoder, data, image.size[0], image.size[1], len(sizes)))

    dds.flags = DDSFlags.TEXTURE | DDSFlags.LINEAR_SIZE
    caps0 = Caps0.TEXTURE
    if len(sizes) > 1:
        dds.flags |= DDSFlags.MIPMAPCOUNT
        caps0 |= Caps0.MIPMAP | Caps0.COMPLEX
    dds.caps = (caps0, 0, 0, 0)
    dds.mipmap_count = len(sizes)
    dds.pitch = dds.textures[0].nbytes
    dds.size = dds.textures[0].size
    dds.pf_flags = PFFlags.FOURCC
    dds.four_cc = four_cc
    return dds


def encode_auto(image, level: int = 18, black: bool = False, threecolor: bool = True,
                mip_count: typing.Optional[int] = None) -> DDSFile:
    """``quicktex encode auto`` for one image (quicktex/cli/encode.py:103-141; the
    command's file handling and flip stay with the caller): BC3 (DXT5) when any
    texel's alpha is below 255, else BC1 (DXT1) in the colour mode the flags
    choose.  The alpha test, the mip chain and the encode run on the GPU after one
    upload of the image."""
    mode_enum = bc1.BC1Encoder.ColorMode
    if not threecolor:
        color_mode = mode_enum.FourColor
    elif not black:
        color_mode = mode_enum.ThreeColor
    else:
        color_mode = mode_enum.ThreeColorBlack
    encoder = bc1.BC1Encoder(level, color_mode)

    mode = 'RGBA' if 'A' in image.mode else 'RGBX'
    image.apply_transparency()
    image = image.convert(mode)
    sizes = mip_sizes(image.size, mip_count)
    is_bc3, textures = _native()._encode_mips_auto(encoder, image.tobytes('raw', mode), image.size[0], image.size[1],
                                                   len(sizes))
    dds = DDSFile()
    dds.textures = list(textures)
    dds.flags = DDSFlags.TEXTURE | DDSFlags.LINEAR_SIZE
    caps0 = Caps0.TEXTURE
    if len(sizes) > 1:
        dds.flags |= DDSFlags.MIPMAPCOUNT
        caps0 |= Caps0.MIPMAP | Caps0.COMPLEX
    dds.caps = (caps0, 0, 0, 0)
    dds.mipmap_count = len(sizes)
    dds.pitch = dds.textures[0].nbytes
    dds.size = dds.textures[0].size
    dds.pf_flags = PFFlags.FOURCC
    dds.four_cc = 'DXT5' if is_bc3 else 'DXT1'
    dds.format = dds_formats[1] if is_bc3 else dds_formats[0]
    return dds


def swizzle_bc3nm(image):
    """The normal-map swizzle of ``quicktex encode bc3nm`` (quicktex/cli/encode.py:52-55)
    on the GPU: an RGBA image (255, band 1, band 1, band 0) of an image of two or
    more bands."""
    import numpy as np
    from PIL import Image
    bands = image.split()
    w, h = image.size
    packed = np.zeros((h, w, 4), np.uint8)
    packed[..., 0] = np.asarray(bands[0])
    packed[..., 1] = np.asarray(bands[1])
    out = _native()._swizzle_bc3nm(packed.reshape(-1), w, h)
    return Image.frombytes('RGBA', (w, h), out)
