"""GPU: quicktex_amd.dds.encode and DDSFile.decode_all against the reference's
per-level flow (quicktex/dds.py:276-287, 209-216): Pillow's
resize_no_premultiply of the base for each mip level, then one encoder.encode
per level; one decoder.decode per level.  Bit-exact for every level, on a real
RGBA image (Bun.png from the reference's tests) and synthetic RGB / L images.
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

Image = pytest.importorskip('PIL.Image')

IMAGES = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden', 'images')


@pytest.fixture(scope='module')
def mods():
    import quicktex_amd
    from quicktex_amd import dds, image_utils
    from quicktex_amd.s3tc import bc1, bc3, bc4, bc5, interpolator
    quicktex_amd.interpolator = interpolator
    return quicktex_amd, dds, image_utils, bc1, bc3, bc4, bc5


def _per_level(mods, image, encoder, mip_count=None):
    """The reference's loop, with this package's plain encoder."""
    qt, dds, iu = mods[0], mods[1], mods[2]
    mode = 'RGBA' if 'A' in image.mode else 'RGBX'
    image = image.convert(mode)
    sizes = iu.mip_sizes(image.size, mip_count)
    images = [image] + [iu.resize_no_premultiply(image, size) for size in sizes[1:]]
    return [encoder.encode(qt.RawTexture.frombytes(i.tobytes('raw', mode), *i.size)).tobytes() for i in images]


def _synthetic(mode, w, h, seed):
    rng = np.random.default_rng(seed)
    yy, xx = np.mgrid[0:h, 0:w]
    base = (np.sin(xx / 17.0) * 60 + np.cos(yy / 11.0) * 60 + 128).astype(np.int32)
    chans = {'RGB': 3, 'L': 1, 'RGBA': 4}[mode]
    arr = np.stack([np.clip(base + rng.integers(-20, 20, (h, w)) + 30 * c, 0, 255) for c in range(chans)], -1)
    arr = arr.astype(np.uint8)
    return Image.fromarray(arr[..., 0] if chans == 1 else arr, mode)


def _encoders(mods):
    bc1, bc3, bc4, bc5 = mods[3:]
    return {
        'bc1_l5': (bc1.BC1Encoder(5), 'DXT1'),
        'bc1_l18_3cb_nvidia': (bc1.BC1Encoder(18, bc1.BC1Encoder.ColorMode.ThreeColorBlack,
                                              mods[0].interpolator.InterpolatorNvidia()), 'DXT1'),
        'bc3_l10': (bc3.BC3Encoder(10), 'DXT5'),
        'bc4': (bc4.BC4Encoder(0), 'ATI1'),
        'bc5': (bc5.BC5Encoder(0, 1), 'ATI2'),
    }


@pytest.mark.parametrize('enc', ['bc1_l5', 'bc1_l18_3cb_nvidia', 'bc3_l10', 'bc4', 'bc5'])
def test_encode_bun(mods, enc):
    dds = mods[1]
    image = Image.open(os.path.join(IMAGES, 'Bun.png'))
    assert image.mode == 'RGBA' and image.size == (512, 622)
    encoder, four_cc = _encoders(mods)[enc]
    f = dds.encode(image.copy(), encoder, four_cc)
    want = _per_level(mods, image.copy(), encoder)
    assert len(f.textures) == len(want) == 10
    for m, (t, w) in enumerate(zip(f.textures, want)):
        assert t.tobytes() == w, (enc, m)
    assert f.mipmap_count == 10 and f.four_cc == four_cc
    assert f.flags == dds.DDSFlags.TEXTURE | dds.DDSFlags.LINEAR_SIZE | dds.DDSFlags.MIPMAPCOUNT
    assert f.caps[0] == dds.Caps0.TEXTURE | dds.Caps0.MIPMAP | dds.Caps0.COMPLEX
    assert f.pitch == f.textures[0].nbytes and f.size == (512, 622)


@pytest.mark.parametrize('mode,w,h', [('RGB', 300, 200), ('L', 129, 65), ('RGBA', 1, 37), ('RGB', 1024, 8)])
@pytest.mark.parametrize('mip_count', [None, 2])
def test_encode_modes(mods, mode, w, h, mip_count):
    dds, bc1 = mods[1], mods[3]
    image = _synthetic(mode, w, h, w * h)
    encoder = bc1.BC1Encoder(10)
    f = dds.encode(image.copy(), encoder, 'DXT1', mip_count)
    want = _per_level(mods, image.copy(), encoder, mip_count)
    assert [t.tobytes() for t in f.textures] == want
    if mip_count == 2:
        assert f.mipmap_count == 2


def test_save_read_decode_all(mods, tmp_path):
    dds, bc3 = mods[1], mods[4]
    image = Image.open(os.path.join(IMAGES, 'Bun.png'))
    f = dds.encode(image, bc3.BC3Encoder(5), 'DXT5')
    p = tmp_path / 'bun.dds'
    f.save(p)
    g = dds.read(p)
    assert g.format.name == 'BC3' and g.mipmap_count == 10 and g.size == (512, 622)
    assert [t.tobytes() for t in g.textures] == [t.tobytes() for t in f.textures]
    all_levels = g.decode_all()
    assert len(all_levels) == 10
    for m, im in enumerate(all_levels):
        one = g.decode(m)
        assert im.size == one.size and im.tobytes() == one.tobytes(), m


@pytest.mark.parametrize('name,args', [('bc1/3color.dds', ()), ('bc1/3color_black.dds', (True,)),
                                       ('bc4/6value.dds', (0,)), ('bc4/8value.dds', (0,))])
def test_decode_fixture_files(mods, oracle, name, args):
    dds = mods[1]
    f = dds.read(os.path.join(IMAGES, name))
    a = f.decode_all(*args)[0]
    b = f.decode(0, *args)
    assert a.tobytes() == b.tobytes()
    got = np.asarray(a)
    blocks = np.frombuffer(f.textures[0].tobytes(), np.uint8)
    if name.startswith('bc4'):
        want = oracle.bc4_decode(blocks, 4, 4, 0)
    else:
        wa = bool(args and args[0])
        want = oracle.bc1_decode(blocks, 4, 4, 0, wa)
        # these two files hold the reference's known-answer blocks; the fixture PNGs
        # are their colours with the midpoint rounded up (the AMD interpolator's 128,
        # where Ideal gives 127)
        amd = f.decode_all(wa, mods[0].interpolator.InterpolatorAMD())[0]
        png = np.asarray(Image.open(os.path.join(IMAGES, name[:-4] + '.png')).convert('RGB'))
        assert (np.asarray(amd)[..., :3] == png).all()
        assert (np.asarray(amd) == oracle.bc1_decode(blocks, 4, 4, 3, wa)).all()
    assert (got == want).all()


def test_decode_all_empty_and_mixed_sizes(mods):
    qt, dds, bc1 = mods[0], mods[1], mods[3]
    f = dds.DDSFile()
    f.format = dds.dds_formats[0]
    assert f.decode_all() == []
    enc = bc1.BC1Encoder(3)
    rng = np.random.default_rng(1)
    for w, h in ((37, 5), (1, 1), (64, 64)):
        px = rng.integers(0, 256, w * h * 4, dtype=np.uint8)
        f.textures.append(enc.encode(qt.RawTexture.frombytes(px, w, h)))
    dec = f.decode_all(False)
    for t, im in zip(f.textures, dec):
        assert im.tobytes() == bc1.BC1Decoder(False).decode(t).tobytes()
