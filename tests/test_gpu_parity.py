"""GPU parity: the HIP kernels, called through the C ABI, against the CPU checker.

The checker is the reference's own C++ (oracle/_ref) where it was built, else
the oracle restatement.  Bit-exact for every byte: BC1 at every level (0-19),
colour mode and interpolator on a mixture image and a stress image of
two/three-colour, greyscale and near-black blocks, through each of the two BC1
encoders (one block per lane; 16 lanes per block); BC3 with varying alpha;
BC4/BC5; every decoder, on ragged sizes.

Level 1 (EndpointMode.LeastSquares) is compared with the reference built with
zero-initialised automatic variables (oracle/_ref/libqtxref_zi.so): the
reference accumulates into an uninitialised array there
(quicktex/s3tc/bc1/BC1Encoder.cpp:510-515), so its g++ build's output depends on
stack contents; the kernels start that accumulator at zero.
"""
import numpy as np
import pytest

from quicktex_amd import synth
from tests._images import stress_rgba

pytestmark = pytest.mark.gpu

LEVELS = list(range(20))
MODES = [4, 3, 0x13]
INTERPS = [0, 1, 2, 3]


def _mismatch(a, b, bs):
    a = a.reshape(-1, bs)
    b = b.reshape(-1, bs)
    bad = np.nonzero((a != b).any(axis=1))[0]
    return bad


@pytest.fixture(scope='module')
def images():
    return [synth.mixture_rgba(512, 32, seed=0x5EED0000 + 7, near_black=0.15), stress_rgba(256, 32, 11)]


ENCODERS = {'lane': 1, 'group': 2}  # QTX_BC1_LANE, QTX_BC1_GROUP: both encoders, identical bytes


@pytest.fixture(params=sorted(ENCODERS))
def encoder(qtx, request):
    """Runs the test once with each BC1/BC3 encoder forced on the context."""
    qtx.set_bc1_encoder(ENCODERS[request.param])
    yield request.param
    qtx.set_bc1_encoder(0)


@pytest.fixture(scope='module')
def checker():
    from tests._cpu_libs import oracle, reference
    return reference() or oracle()


def checker_for(p, oracle):
    """The compiled reference that defines the bytes for settings p (the
    zero-initialising build for the least-squares endpoint mode), else the oracle."""
    from tests._cpu_libs import checker_for as _c
    return _c(p) or oracle


@pytest.mark.parametrize('interp', INTERPS)
@pytest.mark.parametrize('mode', MODES)
@pytest.mark.parametrize('level', LEVELS)
def test_bc1_encode_levels(qtx, oracle, checker, images, level, mode, interp, encoder):
    p = qtx.params(level, mode, interp)
    assert p.as_tuple() == oracle.params(level, mode, interp).as_tuple()
    cpu = checker_for(p, oracle)
    for k, img in enumerate(images):
        got = qtx.bc1_encode(img, p)
        want = cpu.bc1_encode(img, p)
        bad = _mismatch(got, want, 8)
        assert len(bad) == 0, 'image %d blocks %s: gpu %s cpu %s' % (
            k, bad[:4].tolist(), got.reshape(-1, 8)[bad[0]].tobytes().hex(), want.reshape(-1, 8)[bad[0]].tobytes().hex())


@pytest.mark.parametrize('interp', INTERPS)
@pytest.mark.parametrize('level', [0, 2, 5, 10, 13, 18])
def test_bc3_encode_alpha(qtx, checker, level, interp, encoder):
    img = synth.mixture_rgba(48, 40, seed=99)
    rng = np.random.default_rng(level * 4 + interp)
    img[..., 3] = rng.integers(0, 256, img.shape[:2], dtype=np.uint8)
    img[:8, :8, 3] = 255  # some opaque blocks
    p = qtx.params(level, 4, interp)
    got = qtx.bc3_encode(img, p)
    want = checker.bc3_encode(img, p)
    bad = _mismatch(got, want, 16)
    assert len(bad) == 0, 'blocks %s' % bad[:8].tolist()


@pytest.mark.parametrize('w,h', [(1, 1), (2, 3), (3, 5), (5, 5), (7, 9), (37, 23), (64, 4)])
def test_bc1_encode_ragged(qtx, oracle, w, h, encoder):
    img = synth.mixture_rgba(w, h, seed=w * 100 + h, near_black=0.1)
    for level in (3, 18):
        p = qtx.params(level, 0x13, 0)
        assert (qtx.bc1_encode(img, p) == oracle.bc1_encode(img, p)).all(), (w, h, level)


@pytest.mark.parametrize('w,h', [(1, 1), (2, 3), (5, 5), (9, 7), (123, 77), (256, 256)])
def test_bc4_bc5_encode(qtx, oracle, w, h):
    img = synth.smooth_channels(w, h, seed=w + 3 * h)
    rng = np.random.default_rng(w * h)
    img[..., 2:] = rng.integers(0, 256, (h, w, 2), dtype=np.uint8)
    for ch in range(4):
        assert (qtx.bc4_encode(img, ch) == oracle.bc4_encode(img, ch)).all(), ch
    for c0, c1 in ((0, 1), (2, 3), (1, 1)):
        assert (qtx.bc5_encode(img, c0, c1) == oracle.bc5_encode(img, c0, c1)).all(), (c0, c1)


@pytest.mark.parametrize('w,h', [(1, 1), (2, 3), (4, 4), (5, 5), (9, 7), (123, 77), (256, 128)])
def test_decoders(qtx, oracle, w, h):
    rng = np.random.default_rng(w * 1000 + h)
    nb = ((w + 3) // 4) * ((h + 3) // 4)
    b8 = rng.integers(0, 256, nb * 8, dtype=np.uint8)
    b16 = rng.integers(0, 256, nb * 16, dtype=np.uint8)
    for interp in INTERPS:
        for wa in (False, True):
            assert (qtx.bc1_decode(b8, w, h, interp, wa) == oracle.bc1_decode(b8, w, h, interp, wa)).all(), (interp, wa)
        assert (qtx.bc3_decode(b16, w, h, interp) == oracle.bc3_decode(b16, w, h, interp)).all(), interp
    for ch in range(4):
        assert (qtx.bc4_decode(b8, w, h, ch) == oracle.bc4_decode(b8, w, h, ch)).all(), ch
    for c0, c1 in ((0, 1), (3, 2), (1, 1)):
        assert (qtx.bc5_decode(b16, w, h, c0, c1) == oracle.bc5_decode(b16, w, h, c0, c1)).all(), (c0, c1)
