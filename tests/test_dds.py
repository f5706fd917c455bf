"""DDS container (quicktex_amd.dds) on the CPU: the reader against DDS files the
reference's tests hold (tests/golden/images, copied by copy_images.py), the
writer's byte layout (Microsoft's DDS_HEADER), and mip_sizes.  Encode and
decode_all run on the GPU (test_gpu_dds.py)."""
import os
import struct

import pytest

IMAGES = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden', 'images')

# block bytes of each fixture: the BC1 ones are the reference's known-answer blocks
# (tests/images/__init__.py:18-20)
FIXTURES = {
    'bc1/3color.dds': ('DXT1', 'e00700f829292929'),
    'bc1/3color_black.dds': ('DXT1', 'e00700f827272727'),
    'bc1/greyscale.dds': ('DXT1', 'ffff494a78787878'),
    'bc4/6value.dds': ('ATI1', 'f708c11bbc544005'),
    'bc4/8value.dds': ('ATI1', 'f00f888668accffa'),
}


@pytest.fixture(scope='module')
def dds():
    pytest.importorskip('PIL')
    try:
        from quicktex import dds as d
    except ImportError as e:
        pytest.skip('native module not built: %s' % e)
    return d


@pytest.mark.parametrize('name', sorted(FIXTURES))
def test_read_fixture(dds, name):
    four_cc, block = FIXTURES[name]
    f = dds.read(os.path.join(IMAGES, name))
    assert f.four_cc == four_cc
    assert f.format.four_cc == four_cc
    assert f.size == (4, 4)
    assert f.mipmap_count == 1 and f.depth == 1  # flags without MIPMAPCOUNT / DEPTH
    assert f.flags == dds.DDSFlags.TEXTURE | dds.DDSFlags.LINEAR_SIZE
    assert f.pitch == 8
    assert len(f.textures) == 1
    assert f.textures[0].tobytes().hex() == block


@pytest.mark.parametrize('name', sorted(FIXTURES))
def test_save_read_round_trip(dds, name, tmp_path):
    src = os.path.join(IMAGES, name)
    f = dds.read(src)
    out = tmp_path / 'copy.dds'
    f.save(out)
    a, b = open(src, 'rb').read(), open(out, 'rb').read()
    assert len(a) == len(b) == 136
    # identical but for depth and mipmap count, which the reader fills with 1
    # when their flags are absent (the fixtures hold 0 there), and the 44 reserved
    # bytes, which the reader skips and the writer zeroes (the BC4 files carry a
    # writer's tag there)
    assert a[:24] == b[:24] and a[76:] == b[76:]
    assert struct.unpack('<2I', b[24:32]) == (1, 1)
    assert b[32:76] == bytes(44)
    g = dds.read(out)
    assert g.textures[0].tobytes() == f.textures[0].tobytes()


def test_header_layout(dds, tmp_path):
    from quicktex.s3tc.bc1 import BC1Texture
    f = dds.DDSFile()
    f.flags = dds.DDSFlags.TEXTURE | dds.DDSFlags.LINEAR_SIZE | dds.DDSFlags.MIPMAPCOUNT
    f.size = (12, 8)
    f.pitch = 48
    f.mipmap_count = 2
    f.four_cc = 'DXT1'
    f.caps = (dds.Caps0.TEXTURE | dds.Caps0.MIPMAP | dds.Caps0.COMPLEX, 0, 0, 0)
    f.textures = [BC1Texture(12, 8), BC1Texture(6, 4)]
    p = tmp_path / 'h.dds'
    f.save(p)
    b = open(p, 'rb').read()
    assert b[:4] == b'DDS '
    # DDS_HEADER: size, flags, height, width, pitchOrLinearSize, depth, mipMapCount, 11 reserved
    assert struct.unpack('<7I', b[4:32]) == (124, 0xA1007, 8, 12, 48, 1, 2)
    assert b[32:76] == bytes(44)
    # DDS_PIXELFORMAT: size, flags, fourCC, RGB bit count, 4 masks
    assert struct.unpack('<2I', b[76:84]) == (32, 0x4) and b[84:88] == b'DXT1'
    assert struct.unpack('<5I', b[88:108]) == (0, 0, 0, 0, 0)
    # caps, caps2-4, reserved
    assert struct.unpack('<4I', b[108:124]) == (0x401008, 0, 0, 0) and b[124:128] == bytes(4)
    assert len(b) == 128 + 48 + 16


def test_mip_sizes():
    from quicktex.image_utils import mip_sizes
    assert mip_sizes((63, 63))[:2] == [(63, 63), (31, 31)]  # the docstring's example (image_utils.py:12)
    assert mip_sizes((4096, 4096))[-1] == (1, 1) and len(mip_sizes((4096, 4096))) == 13
    assert mip_sizes((1000, 600)) == [(1000, 600), (500, 300), (250, 150), (125, 75), (62, 37), (31, 18), (15, 9),
                                      (7, 4), (3, 2), (1, 1)]
    assert mip_sizes((5, 1)) == [(5, 1), (2, 1), (1, 1)]
    assert mip_sizes((8, 8), 2) == [(8, 8), (4, 4)]
    assert mip_sizes((1, 1)) == [(1, 1)]
    with pytest.raises(AssertionError):
        mip_sizes((0, 4))


def test_mip_sizes_definition():
    """Against the chain's definition, level by level: halve with floor, stop at 1,
    end at the first 1 x 1 level or after mip_count levels."""
    from quicktex.image_utils import mip_sizes
    for w in list(range(1, 70)) + [127, 128, 129, 1000, 4095, 4096, 4097, 65535, 65536]:
        for h in (1, 2, 3, 17, 64, 600, w):
            for count in (None, 0, 1, 2, 5, 40):
                chain, size = [], (w, h)
                while count is None or count == 0 or len(chain) < count:
                    chain.append(size)
                    if size == (1, 1):
                        break
                    size = (max(size[0] // 2, 1), max(size[1] // 2, 1))
                assert mip_sizes((w, h), count) == chain, (w, h, count)
    with pytest.raises(AssertionError):
        mip_sizes((4, 4), -1)


def test_reader_errors(dds, tmp_path):
    """The reference reader's exceptions: AssertionError for a bad magic, header or
    pixel-format size and a short file, NotImplementedError for DX10, StopIteration
    for an unknown FourCC."""
    src = open(os.path.join(IMAGES, 'bc1/greyscale.dds'), 'rb').read()
    cases = {
        'magic': (b'DDX ' + src[4:], AssertionError),
        'header': (src[:4] + struct.pack('<I', 120) + src[8:], AssertionError),
        'pixel_format': (src[:76] + struct.pack('<I', 31) + src[80:], AssertionError),
        'short': (src[:-1], AssertionError),
        'dx10': (src[:84] + b'DX10' + src[88:], NotImplementedError),
        'fourcc': (src[:84] + b'DXT3' + src[88:], StopIteration),
    }
    for name, (data, exc) in cases.items():
        path = tmp_path / (name + '.dds')
        path.write_bytes(data)
        with pytest.raises(exc):
            dds.read(path)
