"""Pins of the oracle (oracle/qtx_oracle.cpp), CPU only.

Each part of the restatement is checked against something that does not share
its code: the reference test suite's known answers, vectors written by the
reference's own C++ (tests/golden/ref_vectors.npz), the compiled reference
itself where it was built, and closed forms / brute force.
"""
import json
import os

import numpy as np
import pytest

from quicktex_amd import synth

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden')


@pytest.fixture(scope='module')
def kat():
    return json.load(open(os.path.join(GOLDEN, 'reference_tests.json')))


@pytest.fixture(scope='module')
def vec():
    return np.load(os.path.join(GOLDEN, 'ref_vectors.npz'))


def tex4(texels):
    return np.array(texels, dtype=np.uint8).reshape(4, 4, 4)


# -- the reference test suite's known answers (test_bc1.py:134-206, test_bc4.py:145-179)

@pytest.mark.parametrize('mode', [4, 3, 0x13])
def test_bc1_encoder_goldens(oracle, kat, mode):
    o = oracle
    p = o.params(5, mode, 0)
    g = kat['bc1']
    got = o.bc1_encode(tex4(g['greyscale']['texels']), p).tobytes().hex()
    assert got == g['greyscale']['block']  # test_block_4color
    three = o.bc1_encode(tex4(g['three_color']['texels']), p).tobytes()
    c0, c1 = int.from_bytes(three[0:2], 'little'), int.from_bytes(three[2:4], 'little')
    if mode != 4:
        assert three.hex() == g['three_color']['block']  # test_block_3color
    else:
        assert c0 > c1  # never 3-colour in FourColor mode
    blk = o.bc1_encode(tex4(g['three_color_black']['texels']), p).tobytes()
    c0, c1 = int.from_bytes(blk[0:2], 'little'), int.from_bytes(blk[2:4], 'little')
    sels = [(blk[4 + i // 4] >> (2 * (i % 4))) & 3 for i in range(16)]
    if mode == 0x13:
        assert blk.hex() == g['three_color_black']['block']
        assert c0 <= c1 and 3 in sels
    elif mode == 3:
        assert not (3 in sels and c0 <= c1)
    else:
        assert c0 > c1


def test_bc1_decoder_goldens(oracle, kat):
    for name in ('greyscale', 'three_color', 'three_color_black'):
        g = kat['bc1'][name]
        out = oracle.bc1_decode(np.frombuffer(bytes.fromhex(g['block']), np.uint8), 4, 4)
        assert out.reshape(-1).tolist() == g['texels'], name


def test_bc4_goldens(oracle, kat):
    g = kat['bc4']['eight_value']
    blk = oracle.bc4_encode(tex4(g['texels']), 0).tobytes()
    assert blk.hex() == kat['bc4_block']['bytes']  # test_bc4.py:145-157
    for name in ('six_value', 'eight_value'):
        g = kat['bc4'][name]
        a0, a1 = g['endpoints']
        bits = 0
        for i, s in enumerate(sum(g['selectors'], [])):
            bits |= s << (3 * i)
        block = bytes([a0, a1]) + bits.to_bytes(6, 'little')
        out = oracle.bc4_decode(np.frombuffer(block, np.uint8), 4, 4, 0)
        want = np.array(g['texels'], np.uint8).reshape(4, 4, 4)
        assert (out[..., 0] == want[..., 0]).all(), name  # test_bc4.py:160-179 compares channel 0 via the PNG


# -- vectors written by the reference's compiled C++

def test_bc1_ref_vectors(oracle, vec):
    img = vec['bc1_input']
    n = 0
    for key in vec.files:
        if not key.startswith('bc1_L'):
            continue
        lvl, mode, interp = [int(x[1:]) for x in key.split('_')[1:]]
        got = oracle.bc1_encode(img, oracle.params(lvl, mode, interp))
        assert (got == vec[key]).all(), key
        n += 1
    assert n == 20 * 3 * 4  # level 1 from the zero-initialising reference build


def test_bc345_ref_vectors(oracle, vec):
    a = vec['bc3_input']
    for interp in range(4):
        for level in (0, 5, 10, 18):
            assert (oracle.bc3_encode(a, oracle.params(level, 4, interp)) == vec['bc3_L%d_I%d' % (level, interp)]).all()
    sm = vec['bc45_input']
    for ch in range(4):
        assert (oracle.bc4_encode(sm, ch) == vec['bc4_C%d' % ch]).all()
    assert (oracle.bc5_encode(sm, 0, 1) == vec['bc5_C01']).all()
    assert (oracle.bc5_encode(sm, 3, 2) == vec['bc5_C32']).all()


def test_decoder_ref_vectors(oracle, vec):
    w, h = 37, 23
    b8, b16 = vec['dec_blocks8'], vec['dec_blocks16']
    for interp in range(4):
        for wa in (0, 1):
            assert (oracle.bc1_decode(b8, w, h, interp, bool(wa)) == vec['dec_bc1_I%d_A%d' % (interp, wa)]).all()
        assert (oracle.bc3_decode(b16, w, h, interp) == vec['dec_bc3_I%d' % interp]).all()
    for ch in range(4):
        assert (oracle.bc4_decode(b8, w, h, ch) == vec['dec_bc4_C%d' % ch]).all()
    assert (oracle.bc5_decode(b16, w, h, 0, 1) == vec['dec_bc5_C01']).all()


# -- the compiled reference, where it was built here

@pytest.mark.parametrize('interp', [0, 1, 2, 3])
def test_params_match_reference(oracle, reference, interp):
    for mode in (4, 3, 0x13):
        for level in range(20):
            assert oracle.params(level, mode, interp).as_tuple() == reference.params(level, mode, interp).as_tuple()


@pytest.mark.parametrize('interp', [0, 1, 2, 3])
def test_bc1_matches_reference(oracle, reference, interp):
    img = synth.mixture_rgba(256, 32, seed=0x1234 + interp, near_black=0.2)
    for mode in (4, 3, 0x13):
        for level in (0, 2, 3, 4, 5, 8, 9, 12, 13, 16, 17, 18):
            p = oracle.params(level, mode, interp)
            assert (oracle.bc1_encode(img, p) == reference.bc1_encode(img, p)).all(), (level, mode)


# -- closed forms and brute force

def test_bc4_encode_is_optimal(oracle):
    """Every texel decodes to a palette value nearest to it (BC4Encoder.cpp:43-44)."""
    img = synth.smooth_channels(64, 64, seed=5, nchan=1)
    rng = np.random.default_rng(3)
    img[:32, :, 0] = rng.integers(0, 256, (32, 64), dtype=np.uint8)
    blocks = oracle.bc4_encode(img, 0)
    dec = oracle.bc4_decode(blocks, 64, 64, 0)[..., 0].astype(int)
    for k, blk in enumerate(blocks.reshape(-1, 8)):
        by, bx = divmod(k, 16)
        src = img[by * 4:by * 4 + 4, bx * 4:bx * 4 + 4, 0].astype(int)
        a0, a1 = int(blk[0]), int(blk[1])
        assert a0 >= a1  # always 8-value mode (or solid)
        if a0 == a1:
            pal = [a0]
        else:
            pal = [a0, a1] + [((7 - j) * a0 + j * a1) // 7 for j in range(1, 7)]
        best = np.min(np.abs(src[..., None] - np.array(pal)[None, None, :]), axis=2)
        assert (np.abs(dec[by * 4:by * 4 + 4, bx * 4:bx * 4 + 4] - src) == best).all()


def test_single_color_tables_are_optimal(oracle):
    """Each entry of the match tables attains the minimum error over all endpoint
    pairs under its interpolator (SingleColorTable.h:55-88)."""
    import ctypes
    s5 = lambda v: (v << 3) | (v >> 2)  # noqa: E731
    s6 = lambda v: (v << 2) | (v >> 4)  # noqa: E731
    for interp in range(4):
        for bits in (5, 6):
            buf = (ctypes.c_uint8 * 768)()
            assert oracle.lib.qtxo_match_table(interp, bits, buf) == 0
            t = np.frombuffer(bytes(buf), np.uint8).reshape(256, 3).astype(int)
            n = 1 << bits
            lo, hi = np.meshgrid(np.arange(n), np.arange(n), indexing='ij')
            e = s5 if bits == 5 else s6
            l8, h8 = e(lo), e(hi)
            if interp == 2:  # Nvidia works on the 5/6-bit codes
                if bits == 5:
                    val = ((2 * hi + lo) * 22) // 8
                else:
                    d = lo - hi
                    val = (256 * hi + np.trunc(d / 4).astype(int) + 128 + d * 80) >> 8
            elif interp == 1:
                val = (2 * h8 + l8 + 1) // 3
            elif interp == 3:
                val = (43 * h8 + 21 * l8 + 32) >> 6
            else:
                val = (2 * h8 + l8) // 3
            pen = (np.abs(h8 - l8) * 3) // 100 if interp in (0, 1) else 0
            for i in range(256):
                err = np.abs(val - i) + pen
                assert t[i, 2] == err.min(), (interp, bits, i)
                assert err[t[i, 1], t[i, 0]] == t[i, 2]


def test_midpoints(oracle):
    """Rounding points sit between adjacent expanded codes; where the reference
    tree is present they equal its literal table (quicktex/Color.h:87-95)."""
    import ctypes
    import re
    m5, m6 = (ctypes.c_float * 32)(), (ctypes.c_float * 64)()
    oracle.lib.qtxo_midpoints(m5, m6)
    m5, m6 = np.array(m5[:]), np.array(m6[:])
    for m, bits in ((m5, 5), (m6, 6)):
        n = 1 << bits
        e = [((v << 3) | (v >> 2)) if bits == 5 else ((v << 2) | (v >> 4)) for v in range(n)]
        for i in range(n - 1):
            assert e[i] / 255 < m[i] < e[i + 1] / 255
        assert m[n - 1] == np.float32(1e37)
    src = '/root/reference/quicktex/Color.h'
    if os.path.exists(src):
        text = open(src).read()
        for name, m in (('Midpoints5bit', m5), ('Midpoints6bit', m6)):
            body = re.search(name + r'\[\d+\] = \{([^}]*)\}', text).group(1)
            lit = np.array([np.float32(float(x.strip().rstrip('f'))) for x in body.split(',')], np.float32)
            assert (lit == m.astype(np.float32)).all()


def test_order_factors_invert_normal_matrix(oracle):
    """factors x normal-equation matrix = (d/255) I for every non-singular ordering
    (OrderTable.h:63-82)."""
    import ctypes
    import re
    src = open(os.path.join(os.path.dirname(GOLDEN), '..', 'quicktex_amd', 'csrc', 'order_tables.inc')).read()

    def arr(name):
        return [int(x) for x in re.search(name + r'\[[^\]]*\] = \{([^}]*)\}', src).group(1).replace('\n', '').split(',')
                if x.strip()]
    for n, name in ((4, 'qtx_orders4'), (3, 'qtx_orders3')):
        hist = np.array(arr(name)).reshape(-1, n)
        f = (ctypes.c_float * (4 * len(hist)))()
        oracle.lib.qtxo_order_factors(n, f)
        f = np.array(f[:], np.float64).reshape(-1, 4)
        d = n - 1
        for h, fm in zip(hist, f):
            s = np.arange(n)
            m = np.array([[np.sum(h * s * s), np.sum(h * s * (d - s))], [np.sum(h * s * (d - s)), np.sum(h * (d - s) ** 2)]])
            if abs(np.linalg.det(m)) < 1e-8:
                assert (fm == 0).all()
                continue
            prod = fm.reshape(2, 2) @ m
            assert np.allclose(prod, np.eye(2) * d / 255, atol=1e-6)


# -- mip levels: Pillow's BILINEAR resize (quicktex/image_utils.py:42-52) -----------------

def _pil_resize_no_premultiply(rgba, ow, oh):
    """quicktex.image_utils.resize_no_premultiply on an RGBA image, through Pillow itself."""
    from PIL import Image
    im = Image.frombytes('RGBA', (rgba.shape[1], rgba.shape[0]), np.ascontiguousarray(rgba).tobytes())
    rgb = im.convert('RGB').resize((ow, oh), Image.BILINEAR)
    rgb.putalpha(im.getchannel('A').resize((ow, oh), Image.BILINEAR))
    return np.asarray(rgb)


def _chain(w, h):
    out = []
    while (w, h) != (1, 1):
        w, h = max(w // 2, 1), max(h // 2, 1)
        out.append((w, h))
    return out


@pytest.mark.parametrize('w,h', [(1, 2), (2, 1), (3, 3), (5, 7), (13, 1), (1, 29), (37, 23), (64, 64),
                                 (255, 3), (256, 256), (1000, 600), (513, 257)])
def test_resize_matches_pillow(oracle, w, h):
    pytest.importorskip('PIL')
    rng = np.random.default_rng(w * 7919 + h)
    img = rng.integers(0, 256, (h, w, 4), dtype=np.uint8)
    for ow, oh in _chain(w, h) + [(max(1, w * 2 // 3), max(1, h * 3 // 5)), (w + 3, h + 1)]:
        got = oracle.resize(img, ow, oh)
        want = _pil_resize_no_premultiply(img, ow, oh)
        assert (got == want).all(), (w, h, ow, oh)


def test_resize_rgbx_equals_per_channel(oracle):
    """Image.resize on RGBX (dds.encode of an image without alpha) resamples the
    fourth band like the others: the same as the per-channel restatement."""
    from PIL import Image
    rng = np.random.default_rng(5)
    img = rng.integers(0, 256, (45, 70, 4), dtype=np.uint8)
    img[..., 3] = 255
    im = Image.frombytes('RGBX', (70, 45), img.tobytes())
    for ow, oh in _chain(70, 45):
        want = np.frombuffer(im.resize((ow, oh), Image.BILINEAR).tobytes('raw', 'RGBX'), np.uint8).reshape(oh, ow, 4)
        assert (oracle.resize(img, ow, oh) == want).all(), (ow, oh)


def test_resize_closed_forms(oracle):
    """Constant images stay constant; a 2:1 shrink of a 2-pixel row is the rounded mean."""
    img = np.full((9, 11, 4), 77, np.uint8)
    assert (oracle.resize(img, 5, 4) == 77).all()
    row = np.array([[[10, 20, 30, 40], [11, 25, 32, 200]]], np.uint8)
    want = (row.astype(int).sum(1) + 1) // 2  # weights 1/2, 1/2 -> (a + b + 1) >> 1
    assert (oracle.resize(row, 1, 1)[0, 0] == want[0]).all()
