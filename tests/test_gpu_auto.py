"""GPU: the per-image choices of the reference's `quicktex encode` commands
(quicktex/cli/encode.py): the alpha test of `encode auto` against numpy, the
auto-routed chain encode against the reference's flow (Pillow histogram test,
then dds.encode with BC3Encoder(level) or BC1Encoder(level, mode)), and the
`bc3nm` swizzle against Pillow's split/merge."""
import os

import numpy as np
import pytest

from quicktex_amd.capi import FMT_BC1, FMT_BC3, has_alpha_host, mips_encode_auto_host, swizzle_bc3nm_host

pytestmark = pytest.mark.gpu

Image = pytest.importorskip('PIL.Image')
IMAGES = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden', 'images')


@pytest.mark.parametrize('w,h', [(1, 1), (3, 1), (5, 3), (64, 64), (1023, 7), (4096, 1)])
def test_has_alpha(qtx, w, h):
    rng = np.random.default_rng(w * h)
    img = rng.integers(0, 256, (h, w, 4), dtype=np.uint8)
    img[..., 3] = 255
    assert has_alpha_host(qtx, img) is False
    for pos in {0, w * h // 2, w * h - 1}:
        t = img.copy().reshape(-1, 4)
        t[pos, 3] = rng.integers(0, 255)
        assert has_alpha_host(qtx, t.reshape(h, w, 4)) is True, pos
    img[..., 3] = rng.integers(0, 256, (h, w), dtype=np.uint8)
    assert has_alpha_host(qtx, img) == bool((img[..., 3] < 255).any())


def _reference_auto(mods, image, level, black, threecolor):
    """cli/encode.py:117-138 for one image, with this package's encoders."""
    dds, bc1, bc3 = mods
    mode_enum = bc1.BC1Encoder.ColorMode
    mode = mode_enum.FourColor if not threecolor else (mode_enum.ThreeColor if not black else mode_enum.ThreeColorBlack)
    if 'A' not in image.mode:
        has_alpha = False
    else:
        has_alpha = any([a > 0 for a in image.getchannel('A').histogram()[:-1]])
    if has_alpha:
        return 'DXT5', dds.encode(image, bc3.BC3Encoder(level), 'DXT5')
    return 'DXT1', dds.encode(image, bc1.BC1Encoder(level, mode), 'DXT1')


@pytest.fixture(scope='module')
def mods():
    from quicktex_amd import dds
    from quicktex_amd.s3tc import bc1, bc3
    return dds, bc1, bc3


@pytest.mark.parametrize('case', ['bun', 'bun_opaque', 'rgb', 'la'])
@pytest.mark.parametrize('flags', [(10, False, True), (5, True, True), (8, False, False)])
def test_encode_auto_matches_reference_flow(mods, case, flags):
    dds = mods[0]
    level, black, threecolor = flags
    if case.startswith('bun'):
        image = Image.open(os.path.join(IMAGES, 'Bun.png'))
        if case == 'bun_opaque':
            arr = np.asarray(image.convert('RGBA')).copy()
            arr[..., 3] = 255
            image = Image.fromarray(arr, 'RGBA')
    elif case == 'rgb':
        rng = np.random.default_rng(3)
        image = Image.fromarray(rng.integers(0, 256, (77, 130, 3), dtype=np.uint8), 'RGB')
    else:
        rng = np.random.default_rng(4)
        image = Image.fromarray(rng.integers(0, 256, (64, 40, 2), dtype=np.uint8), 'LA')
    four_cc, want = _reference_auto(mods, image.copy(), level, black, threecolor)
    got = dds.encode_auto(image.copy(), level, black, threecolor)
    assert got.four_cc == four_cc
    assert [t.tobytes() for t in got.textures] == [t.tobytes() for t in want.textures]


def test_encode_auto_capi(qtx):
    rng = np.random.default_rng(8)
    img = rng.integers(0, 256, (40, 60, 4), dtype=np.uint8)
    img[..., 3] = 255
    p = qtx.params(7, 3, 0)
    fmt, levels = mips_encode_auto_host(qtx, p, img, 3)
    assert fmt == FMT_BC1 and len(levels) == 3 and levels[0].size == 15 * 10 * 8
    img[5, 5, 3] = 7
    fmt, levels = mips_encode_auto_host(qtx, p, img, 3)
    assert fmt == FMT_BC3 and levels[0].size == 15 * 10 * 16
    # the BC3 colour blocks use the params in FourColor
    p4 = qtx.params(7, 4, 0)
    assert (levels[0] == qtx.bc3_encode(img, p4)).all()


@pytest.mark.parametrize('mode', ['RGB', 'RGBA', 'LA'])
def test_swizzle_bc3nm(mods, mode):
    dds = mods[0]
    rng = np.random.default_rng(len(mode))
    chans = {'RGB': 3, 'RGBA': 4, 'LA': 2}[mode]
    image = Image.fromarray(rng.integers(0, 256, (33, 47, chans), dtype=np.uint8), mode)
    bands = image.split()
    one = Image.new('L', image.size, 0xFF)
    want = Image.merge('RGBA', (one, bands[1], bands[1], bands[0]))  # cli/encode.py:52-55
    got = dds.swizzle_bc3nm(image)
    assert got.mode == 'RGBA' and got.tobytes() == want.tobytes()


def test_swizzle_capi(qtx):
    rng = np.random.default_rng(9)
    img = rng.integers(0, 256, (5, 7, 4), dtype=np.uint8)
    got = swizzle_bc3nm_host(qtx, img)
    want = np.stack([np.full((5, 7), 255, np.uint8), img[..., 1], img[..., 1], img[..., 0]], -1)
    assert (got == want).all()
