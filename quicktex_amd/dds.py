"""DDS files of S3TC textures, with the names and file format of quicktex.dds and
the mip chain built and encoded on the GPU.

The container is Microsoft's DDS: the magic ``DDS ``, a 124-byte DDS_HEADER
holding a 32-byte DDS_PIXELFORMAT, then every mip level's blocks, largest
first.  The header is described once, by the field table ``HEADER_FIELDS``,
and read and written with one ``struct`` layout built from it.  The public
names, flag values, FourCC codes, defaults and exceptions are those of the
reference module (quicktex/dds.py:17-303), so files and callers carry over:

* :func:`encode` converts the image with Pillow as the reference does
  (``apply_transparency``, then RGBA or RGBX: quicktex/dds.py:276-279), then
  resamples every mip level from the base and encodes all of them in one
  batched GPU launch (include/qtx.h, qtx_mips_encode_host); the levels equal
  Pillow's BILINEAR resize and the blocks equal per-level ``encoder.encode``.
* :meth:`DDSFile.decode_all` decodes every level in one batched launch.
* :func:`encode_auto` and :func:`swizzle_bc3nm` are the per-image choices of the
  reference's ``quicktex encode auto`` / ``encode bc3nm`` commands
  (quicktex/cli/encode.py:52-55, 103-141), with the alpha test and the swizzle
  on the GPU.
"""
from __future__ import annotations

import dataclasses
import enum
import os
import struct
import typing

from . import _native
from .image_utils import mip_sizes
from .s3tc import bc1, bc3, bc4, bc5


@dataclasses.dataclass
class DDSFormat:
    """One block format a DDS file can hold: its texture, encoder and decoder
    classes and the FourCC code that names it in the header."""
    name: str
    texture: typing.Any
    encoder: typing.Any
    decoder: typing.Any
    four_cc: typing.Optional[str] = None


def _format(name: str, module, four_cc: str) -> DDSFormat:
    return DDSFormat(name, getattr(module, name + 'Texture'), getattr(module, name + 'Encoder'),
                     getattr(module, name + 'Decoder'), four_cc)


# the four formats and their FourCC codes (DXT1 = BC1, DXT5 = BC3, ATI1 = BC4, ATI2 = BC5)
dds_formats = [_format('BC1', bc1, 'DXT1'), _format('BC3', bc3, 'DXT5'), _format('BC4', bc4, 'ATI1'),
               _format('BC5', bc5, 'ATI2')]


class PFFlags(enum.IntFlag):
    """DDS_PIXELFORMAT.dwFlags: which pixel-format members are valid."""
    ALPHAPIXELS = 0x1   # uncompressed with alpha; the alpha mask is valid
    ALPHA = 0x2         # alpha-only uncompressed data (older files)
    FOURCC = 0x4        # compressed; the FourCC code is valid
    RGB = 0x40          # uncompressed RGB; bit count and colour masks are valid
    YUV = 0x200         # uncompressed YUV (older files)
    LUMINANCE = 0x20000  # single-channel uncompressed (older files)


class DDSFlags(enum.IntFlag):
    """DDS_HEADER.dwFlags: which header members are valid."""
    CAPS = 0x1
    HEIGHT = 0x2
    WIDTH = 0x4
    PITCH = 0x8            # pitch of an uncompressed texture
    PIXEL_FORMAT = 0x1000
    MIPMAPCOUNT = 0x20000
    LINEAR_SIZE = 0x80000  # byte size of a compressed top level
    DEPTH = 0x800000
    TEXTURE = CAPS | HEIGHT | WIDTH | PIXEL_FORMAT  # the members every file must have


class Caps0(enum.IntFlag):
    """DDS_HEADER.dwCaps: surface complexity."""
    COMPLEX = 0x8      # more than one surface (e.g. mip levels)
    MIPMAP = 0x400000
    TEXTURE = 0x1000   # always set


# DDS_HEADER in file order after the magic: (member, struct code); None marks
# reserved bytes, written as zeros and skipped on reading.
HEADER_FIELDS = (
    ('header_bytes', 'I'), ('flags', 'I'), ('height', 'I'), ('width', 'I'), ('pitch', 'I'), ('depth', 'I'),
    ('mipmap_count', 'I'), (None, '44x'),
    # DDS_PIXELFORMAT
    ('pf_bytes', 'I'), ('pf_flags', 'I'), ('four_cc', '4s'), ('pixel_size', 'I'),
    ('mask_r', 'I'), ('mask_g', 'I'), ('mask_b', 'I'), ('mask_a', 'I'),
    ('caps0', 'I'), ('caps1', 'I'), ('caps2', 'I'), ('caps3', 'I'), (None, '4x'),
)
_LAYOUT = struct.Struct('<' + ''.join(code for _, code in HEADER_FIELDS))
_MEMBERS = tuple(name for name, _ in HEADER_FIELDS if name is not None)
PIXEL_FORMAT_BYTES = 32


@typing.final
class DDSFile:
    """A DDS file in memory: the header members and one block texture per mip level.

    Microsoft's reference for the format:
    https://docs.microsoft.com/en-us/windows/win32/direct3ddds/dx-graphics-dds-reference
    """

    magic = b'DDS '
    extension = 'dds'
    header_bytes = 124

    def __init__(self):
        self.flags: DDSFlags = DDSFlags.TEXTURE
        self.size: typing.Tuple[int, int] = (0, 0)  # width, height of level 0
        self.pitch: int = 0                          # compressed: bytes of level 0
        self.depth: int = 1
        self.mipmap_count: int = 1
        self.pf_flags: PFFlags = PFFlags.FOURCC
        self.four_cc: str = 'NONE'
        self.pixel_size: int = 0                     # bits per pixel (uncompressed only)
        self.pixel_bitmasks: typing.Tuple[int, int, int, int] = (0, 0, 0, 0)
        self.caps: typing.Tuple[Caps0, int, int, int] = (Caps0.TEXTURE, 0, 0, 0)
        self.textures: typing.List = []              # block textures, level 0 first
        self.format: DDSFormat = DDSFormat('NONE', None, None, None)

    def header(self) -> bytes:
        """The 124 header bytes of this file."""
        values = {
            'header_bytes': DDSFile.header_bytes, 'flags': int(self.flags), 'width': self.size[0],
            'height': self.size[1], 'pitch': self.pitch, 'depth': self.depth, 'mipmap_count': self.mipmap_count,
            'pf_bytes': PIXEL_FORMAT_BYTES, 'pf_flags': int(self.pf_flags), 'four_cc': self.four_cc.encode('ascii'),
            'pixel_size': self.pixel_size,
        }
        values.update(zip(('mask_r', 'mask_g', 'mask_b', 'mask_a'), self.pixel_bitmasks))
        values.update(zip(('caps0', 'caps1', 'caps2', 'caps3'), (int(c) for c in self.caps)))
        return _LAYOUT.pack(*(values[m] for m in _MEMBERS))

    def save(self, path: os.PathLike) -> None:
        """Writes the magic, the header and every level's blocks to ``path``."""
        head = self.header()
        if len(head) != DDSFile.header_bytes:
            raise AssertionError('error writing file: incorrect header size')
        with open(path, 'wb') as file:
            file.write(DDSFile.magic + head)
            for texture in self.textures:
                file.write(texture)

    def decode(self, mip: int = 0, *args, **kwargs):
        """Level ``mip`` as a Pillow RGBA image; the other arguments construct the
        format's decoder (e.g. ``write_alpha`` for BC1)."""
        from PIL import Image
        raw = self.format.decoder(*args, **kwargs).decode(self.textures[mip])
        return Image.frombuffer('RGBA', raw.size, raw)

    def decode_all(self, *args, **kwargs):
        """Every level as Pillow RGBA images, largest first, decoded in one batched launch."""
        from PIL import Image
        decoder = self.format.decoder(*args, **kwargs)
        return [Image.frombuffer('RGBA', raw.size, raw) for raw in _native()._decode_all(decoder, list(self.textures))]


def _format_of(four_cc: str) -> DDSFormat:
    for fmt in dds_formats:
        if fmt.four_cc == four_cc:
            return fmt
    raise StopIteration('no DDS format with FourCC %r' % four_cc)  # the reference's next() over the table


def read(path: os.PathLike) -> DDSFile:
    """Parses a DDS file holding BC1, BC3, BC4 or BC5 blocks (quicktex/dds.py:219-272):
    AssertionError for a bad magic, header size, pixel-format size or a short
    file, NotImplementedError for DX10 extended headers, StopIteration for any
    other FourCC."""
    with open(path, 'rb') as file:
        if file.read(len(DDSFile.magic)) != DDSFile.magic:
            raise AssertionError('Incorrect magic bytes in DDS file.')
        head = file.read(DDSFile.header_bytes)
        if len(head) >= 4 and struct.unpack_from('<I', head)[0] != DDSFile.header_bytes:
            raise AssertionError('Incorrect DDS header size.')
        if len(head) != DDSFile.header_bytes:
            raise AssertionError('Unexpected EOF')
        h = dict(zip(_MEMBERS, _LAYOUT.unpack(head)))
        if h['pf_bytes'] != PIXEL_FORMAT_BYTES:
            raise AssertionError('Incorrect pixel format size.')

        dds = DDSFile()
        dds.flags = DDSFlags(h['flags'])
        dds.size = (h['width'], h['height'])
        dds.pitch = h['pitch']
        # members whose flag is absent take their defaults
        dds.depth = h['depth'] if DDSFlags.DEPTH in dds.flags else 1
        dds.mipmap_count = h['mipmap_count'] if DDSFlags.MIPMAPCOUNT in dds.flags else 1
        dds.pf_flags = PFFlags(h['pf_flags'])
        dds.four_cc = h['four_cc'].decode() if PFFlags.FOURCC in dds.pf_flags else 'NONE'
        dds.pixel_size = h['pixel_size']
        # the reference reads the four channel masks but keeps its defaults (0, 0, 0, 0)
        # (quicktex/dds.py:241), so a file saved after reading carries zero masks
        dds.caps = (h['caps0'], h['caps1'], h['caps2'], h['caps3'])
        if dds.four_cc == 'DX10':
            raise NotImplementedError('DX10 headers are not yet supported')
        dds.format = _format_of(dds.four_cc)

        for size in mip_sizes(dds.size, dds.mipmap_count):
            level = dds.format.texture(*size)
            if file.readinto(level) != level.nbytes:
                raise AssertionError('Unexpected end of file')
            dds.textures.append(level)
        return dds


def _rgba_bytes(image):
    """The image as the encoder takes it: converted by Pillow to RGBA when it has an
    alpha band, else RGBX, after applying any palette transparency."""
    mode = 'RGBA' if 'A' in image.mode else 'RGBX'
    image.apply_transparency()
    image = image.convert(mode)
    return image.tobytes('raw', mode), image.size


def _container(textures, four_cc: str) -> DDSFile:
    """A DDSFile holding the encoded levels, with the header members the
    reference's writer sets (quicktex/dds.py:289-303)."""
    dds = DDSFile()
    dds.textures = list(textures)
    levels = len(dds.textures)
    dds.flags = DDSFlags.TEXTURE | DDSFlags.LINEAR_SIZE | (DDSFlags.MIPMAPCOUNT if levels > 1 else 0)
    dds.caps = (Caps0.TEXTURE | (Caps0.MIPMAP | Caps0.COMPLEX if levels > 1 else 0), 0, 0, 0)
    dds.mipmap_count = levels
    dds.pitch = dds.textures[0].nbytes
    dds.size = dds.textures[0].size
    dds.pf_flags = PFFlags.FOURCC
    dds.four_cc = four_cc
    return dds


def encode(image, encoder, four_cc: str, mip_count: typing.Optional[int] = None) -> DDSFile:
    """A DDSFile of ``image`` and its mip chain (``mip_count`` levels, default all),
    encoded by ``encoder``: every level is resampled from the base and encoded on
    the GPU in one batch."""
    data, (w, h) = _rgba_bytes(image)
    levels = len(mip_sizes((w, h), mip_count))
    return _container(_native()._encode_mips(encoder, data, w, h, levels), four_cc)


def encode_auto(image, level: int = 18, black: bool = False, threecolor: bool = True,
                mip_count: typing.Optional[int] = None) -> DDSFile:
    """``quicktex encode auto`` for one image (quicktex/cli/encode.py:103-141; the
    command's file handling and flip stay with the caller): BC3 (DXT5) when any
    texel's alpha is below 255, else BC1 (DXT1) in the colour mode the flags
    choose.  The alpha test, the mip chain and the encode run on the GPU after one
    upload of the image."""
    modes = bc1.BC1Encoder.ColorMode
    color_mode = modes.FourColor if not threecolor else (modes.ThreeColorBlack if black else modes.ThreeColor)
    data, (w, h) = _rgba_bytes(image)
    levels = len(mip_sizes((w, h), mip_count))
    is_bc3, textures = _native()._encode_mips_auto(bc1.BC1Encoder(level, color_mode), data, w, h, levels)
    dds = _container(textures, 'DXT5' if is_bc3 else 'DXT1')
    dds.format = _format_of(dds.four_cc)
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
