"""The reference's Python API tests (tests/test_bc1.py, test_bc4.py,
test_texture.py), restated and importing the package the way the reference's
tests do (``from quicktex.s3tc.bc1 import ...``, test_bc1.py:6): the ``quicktex``
package here is this build's codec under the reference's names.  Blocks and
textures are host containers and run on the CPU; encoder/decoder tests are
marked gpu."""
import json
import math
import os

import numpy as np
import pytest

from quicktex import RawTexture
from quicktex.s3tc.bc1 import BC1Block, BC1Decoder, BC1Encoder, BC1Texture
from quicktex.s3tc.bc3 import BC3Block, BC3Decoder, BC3Encoder, BC3Texture
from quicktex.s3tc.bc4 import BC4Block, BC4Decoder, BC4Encoder, BC4Texture
from quicktex.s3tc.bc5 import BC5Block, BC5Decoder, BC5Encoder, BC5Texture
from quicktex.s3tc.interpolator import Interpolator, InterpolatorAMD, InterpolatorNvidia, InterpolatorRound

KAT = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden', 'reference_tests.json')))


def texture4(texels):
    return RawTexture.frombytes(bytes(texels), 4, 4)


class TestBC1Block:  # test_bc1.py:15-59
    bb = bytes.fromhex(KAT['bc1_block']['bytes'])

    def test_size(self):
        assert BC1Block.nbytes == 8 and BC1Block.width == 4 and BC1Block.height == 4 and BC1Block.size == (4, 4)

    def test_buffer(self):
        block = BC1Block()
        mv = memoryview(block)
        mv[:] = self.bb
        assert not mv.readonly and mv.c_contiguous and mv.ndim == 1 and mv.nbytes == 8 and mv.format == 'B'
        assert mv.tobytes() == self.bb == block.tobytes()

    def test_constructor(self):
        k = KAT['bc1_block']
        block = BC1Block(*[tuple(e) for e in k['in_endpoints']], k['selectors'])
        assert block.tobytes() == self.bb
        assert block.selectors == k['selectors']
        assert block.endpoints == tuple(tuple(e) for e in k['out_endpoints'])
        assert not block.is_3color

    def test_frombytes_eq(self):
        assert BC1Block.frombytes(self.bb) == BC1Block.frombytes(self.bb)
        assert BC1Block.frombytes(self.bb).tobytes() == self.bb

    def test_selector_bounds(self):
        with pytest.raises(ValueError):
            BC1Block((0, 0, 0), (1, 1, 1), [[4, 0, 0, 0]] * 4)


@pytest.mark.parametrize('w', [7, 8, 9])
@pytest.mark.parametrize('h', [7, 8, 9])
@pytest.mark.parametrize('Tex,Block', [(BC1Texture, BC1Block), (BC4Texture, BC4Block), (BC3Texture, BC3Block),
                                       (BC5Texture, BC5Block)])
class TestBlockTexture:  # test_bc1.py:63-131, test_bc4.py:74-142
    def test_dimensions(self, w, h, Tex, Block):
        tex = Tex(w, h)
        wb, hb = math.ceil(w / 4), math.ceil(h / 4)
        assert tex.nbytes == Block.nbytes * wb * hb and len(tex.tobytes()) == tex.nbytes
        assert (tex.width, tex.height, tex.size) == (w, h, (w, h))
        assert (tex.width_blocks, tex.height_blocks, tex.size_blocks) == (wb, hb, (wb, hb))

    def test_blocks(self, w, h, Tex, Block):
        tex = Tex(w, h)
        blocks = [[Block.frombytes(bytes([x, y] + [0] * (Block.nbytes - 2))) for x in range(tex.width_blocks)]
                  for y in range(tex.height_blocks)]
        for x in range(tex.width_blocks):
            for y in range(tex.height_blocks):
                tex[x, y] = blocks[y][x]
        b = tex.tobytes()
        for x in range(tex.width_blocks):
            for y in range(tex.height_blocks):
                i = (x + y * tex.width_blocks) * Block.nbytes
                assert tex[x, y] == blocks[y][x]
                assert Block.frombytes(b[i:i + Block.nbytes]) == blocks[y][x]

    def test_subscript(self, w, h, Tex, Block):  # the reference's `text_subscript`, collected here
        tex = Tex(w, h)
        assert tex[-1, -1] == tex[tex.width_blocks - 1, tex.height_blocks - 1]
        with pytest.raises(IndexError):
            _ = tex[tex.width_blocks, tex.height_blocks]
        with pytest.raises(IndexError):
            _ = tex[-1 - tex.width_blocks, -1 - tex.height_blocks]

    def test_buffer(self, w, h, Tex, Block):
        tex = Tex(w, h)
        mv = memoryview(tex)
        data = bytes(range(Block.nbytes)) * tex.width_blocks * tex.height_blocks
        mv[:] = data
        assert not mv.readonly and mv.c_contiguous and mv.nbytes == tex.nbytes and mv.format == 'B'
        assert mv.tobytes() == data


class TestBC4Block:  # test_bc4.py:15-70
    k = KAT['bc4_block']
    bb = bytes.fromhex(k['bytes'])

    def test_constructor(self):
        block = BC4Block(*self.k['endpoints'], self.k['selectors'])
        assert block.tobytes() == self.bb
        assert block.selectors == self.k['selectors']
        assert block.endpoints == tuple(self.k['endpoints'])
        assert BC4Block.frombytes(self.bb) == block

    def test_values(self):
        b6 = BC4Block(8, 248, [[0] * 4] * 4)
        assert b6.values == self.k['values_6'] and b6.is_6value
        b8 = BC4Block(240, 16, [[0] * 4] * 4)
        assert b8.values == self.k['values_8'] and not b8.is_6value

    def test_composites(self):
        a = BC4Block.frombytes(self.bb)
        c = BC1Block.frombytes(bytes.fromhex(KAT['bc1_block']['bytes']))
        b3 = BC3Block(a, c)
        assert b3.tobytes() == self.bb + c.tobytes() and b3.alpha_block == a and b3.color_block == c
        b5 = BC5Block(a, BC4Block())
        assert b5.tobytes() == self.bb + bytes(8) and b5.blocks == (a, BC4Block())


class TestRawTexture:  # test_texture.py:10-60 (a synthetic 1024x1024 RGBX buffer instead of Boilerplate.png)
    width = height = 1024
    data = np.random.default_rng(0).integers(0, 256, 1024 * 1024 * 4, dtype=np.uint8).tobytes()

    def test_size(self):
        tex = RawTexture(self.width, self.height)
        assert tex.nbytes == self.width * self.height * 4 and tex.size == (self.width, self.height)

    def test_pixels(self):
        tex = RawTexture(self.width, self.height)
        tex[0, 0] = (69, 13, 12, 0)
        tex[-1, -1] = (19, 142, 93, 44)
        d = tex.tobytes()
        assert tex[0, 0] == (69, 13, 12, 0) and tex[-1, -1] == (19, 142, 93, 44)
        assert tuple(d[0:4]) == (69, 13, 12, 0) and tuple(d[-4:]) == (19, 142, 93, 44)
        with pytest.raises(IndexError):
            _ = tex[self.width, self.height]
        with pytest.raises(IndexError):
            _ = tex[-1 - self.width, -1 - self.height]

    def test_buffer_frombytes(self):
        tex = RawTexture(self.width, self.height)
        mv = memoryview(tex)
        mv[:] = self.data
        assert not mv.readonly and mv.c_contiguous and mv.format == 'B' and mv.tobytes() == self.data
        assert RawTexture.frombytes(self.data, self.width, self.height).tobytes() == self.data

    def test_errors(self):
        with pytest.raises(ValueError):
            RawTexture(0, 4)
        with pytest.raises(RuntimeError):
            RawTexture.frombytes(b'\0' * 10, 4, 4)


def test_encoder_settings():  # bc1/_bindings.cpp:104-164
    e = BC1Encoder()
    assert e.color_mode == BC1Encoder.ColorMode.FourColor and isinstance(e.interpolator, Interpolator)
    assert e.error_mode == BC1Encoder.ErrorMode.Faster and e.orderings == (1, 1)
    e.set_level(18)
    assert e.orderings == (128, 32) and e.search_rounds == 256 and e.two_cf_passes and e.power_iterations == 6
    e.orderings = (1000, 0)
    assert e.orderings == (128, 1)
    e.power_iterations = 2
    assert e.power_iterations == BC1Encoder.min_power_iterations == 4
    with pytest.raises(ValueError):
        e.set_level(20)
    for cls in (InterpolatorRound, InterpolatorNvidia, InterpolatorAMD):
        assert isinstance(BC1Encoder(5, BC1Encoder.ColorMode.ThreeColor, cls()).interpolator, cls)
    assert BC3Encoder(10).bc1_encoder.orderings == (20, 8)
    with pytest.raises(ValueError):
        BC4Encoder(4)
    assert BC5Encoder(2, 3).channels == (2, 3) and BC4Decoder().channel == 3


@pytest.mark.gpu
@pytest.mark.parametrize('mode', ['FourColor', 'ThreeColor', 'ThreeColorBlack'])
def test_bc1_encoder_goldens(mode):  # test_bc1.py:134-185
    cm = getattr(BC1Encoder.ColorMode, mode)
    enc = BC1Encoder(color_mode=cm)
    g = KAT['bc1']
    out = enc.encode(texture4(g['greyscale']['texels']))
    assert out.size_blocks == (1, 1) and not out[0, 0].is_3color
    assert out[0, 0].tobytes().hex() == g['greyscale']['block']
    blk = enc.encode(texture4(g['three_color']['texels']))[0, 0]
    if mode != 'FourColor':
        assert blk.is_3color and blk.tobytes().hex() == g['three_color']['block']
    else:
        assert not blk.is_3color
    blk = enc.encode(texture4(g['three_color_black']['texels']))[0, 0]
    has_black = 3 in [j for row in blk.selectors for j in row]
    if mode == 'ThreeColorBlack':
        assert blk.is_3color and has_black and blk.tobytes().hex() == g['three_color_black']['block']
    elif mode == 'ThreeColor':
        assert not (has_black and blk.is_3color)
    else:
        assert not blk.is_3color


@pytest.mark.gpu
def test_decoders_goldens():  # test_bc1.py:188-206, test_bc4.py:160-179
    for name in ('greyscale', 'three_color', 'three_color_black'):
        g = KAT['bc1'][name]
        tex = BC1Texture(4, 4)
        tex[0, 0] = BC1Block.frombytes(bytes.fromhex(g['block']))
        assert list(BC1Decoder().decode(tex).tobytes()) == g['texels'], name
    g = KAT['bc4']['eight_value']
    out = BC4Encoder(0).encode(texture4(g['texels']))
    assert out[0, 0].tobytes().hex() == KAT['bc4_block']['bytes'] and not out[0, 0].is_6value
    for name in ('six_value', 'eight_value'):
        g = KAT['bc4'][name]
        tex = BC4Texture(4, 4)
        tex[0, 0] = BC4Block(*g['endpoints'], g['selectors'])
        dec = np.frombuffer(BC4Decoder(0).decode(tex).tobytes(), np.uint8).reshape(16, 4)
        assert (dec[:, 0] == np.array(g['texels'], np.uint8).reshape(16, 4)[:, 0]).all()


@pytest.mark.gpu
def test_round_trips_match_oracle(oracle):
    from quicktex_amd import synth
    img = synth.mixture_rgba(37, 29, seed=77, near_black=0.1)
    img[..., 3] = np.arange(img.size // 4, dtype=np.uint8).reshape(img.shape[:2])
    raw = RawTexture.frombytes(img.tobytes(), 37, 29)
    for interp, cls in enumerate((Interpolator, InterpolatorRound, InterpolatorNvidia, InterpolatorAMD)):
        enc = BC1Encoder(12, BC1Encoder.ColorMode.ThreeColorBlack, cls())
        blocks = enc.encode(raw)
        want = oracle.bc1_encode(img, oracle.params(12, 0x13, interp))
        assert blocks.tobytes() == want.tobytes()
        dec = BC1Decoder(True, cls()).decode(blocks)
        assert dec.tobytes() == oracle.bc1_decode(want, 37, 29, interp, True).tobytes()
        b3 = BC3Encoder(7, cls()).encode(raw)
        assert b3.tobytes() == oracle.bc3_encode(img, oracle.params(7, 4, interp)).tobytes()
        assert BC3Decoder(cls()).decode(b3).tobytes() == oracle.bc3_decode(np.frombuffer(b3.tobytes(), np.uint8), 37, 29, interp).tobytes()
    b5 = BC5Encoder(1, 3).encode(raw)
    assert b5.tobytes() == oracle.bc5_encode(img, 1, 3).tobytes()
    assert BC5Decoder(1, 3).decode(b5).tobytes() == oracle.bc5_decode(np.frombuffer(b5.tobytes(), np.uint8), 37, 29, 1, 3).tobytes()


def test_quicktex_names_are_this_codec():
    """``quicktex`` and ``quicktex_amd`` expose the same objects, so code written
    against the reference's module names runs this codec unchanged."""
    import quicktex
    import quicktex.dds
    import quicktex.image_utils
    import quicktex.s3tc.bc1
    import quicktex_amd
    import quicktex_amd.dds
    import quicktex_amd.s3tc.bc1
    assert quicktex.RawTexture is quicktex_amd.RawTexture and quicktex.Texture is quicktex_amd.Texture
    assert quicktex.__version__ == quicktex_amd.__version__
    assert quicktex.s3tc.bc1.BC1Encoder is quicktex_amd.s3tc.bc1.BC1Encoder is BC1Encoder
    assert quicktex.dds.read is quicktex_amd.dds.read and quicktex.dds.DDSFile is quicktex_amd.dds.DDSFile
    assert quicktex.image_utils.mip_sizes((8, 2)) == [(8, 2), (4, 1), (2, 1), (1, 1)]
    for mod in ('bc1', 'bc3', 'bc4', 'bc5', 'interpolator'):
        import importlib
        a = importlib.import_module('quicktex.s3tc.' + mod)
        b = importlib.import_module('quicktex_amd.s3tc.' + mod)
        assert sorted(a.__all__) == sorted(b.__all__) and all(getattr(a, n) is getattr(b, n) for n in b.__all__)
