"""GPU: the static BC1/BC3 lane encoder over its block list.

Settings below the queued kernel's cost (levels 0-12 and other settings without
long searches) run on the static lane kernel; from level 7's cost up, and with a
queue set lent to the launch, a first kernel stores the solid blocks and lists
the others, and the lane kernel encodes the listed blocks one per lane.  The
lists' counts and the encode's read count live in the lent queue words and must
be back at zero when the launch completes.  Checked here, bit-exact against the CPU checker: every
build of the lane kernel (level 0-12 settings, any settings), BC1 and BC3 (solid
blocks with and without alpha), relaunches on one stream (a smaller texture
after a larger one), launches in flight on more streams than there are queue
sets (the later ones run without a list), one stream with more launches than
queue sets, the automatic route at a size that takes the lane kernel, and the
list's edge cases: every block solid, none solid, one block, ragged sizes.
"""
import numpy as np
import pytest

from quicktex_amd import synth

pytestmark = pytest.mark.gpu

LANE = 1  # QTX_BC1_LANE: the lane encoder at every size


@pytest.fixture(scope='module')
def checker():
    from tests._cpu_libs import oracle, reference
    return reference() or oracle()


@pytest.fixture
def lane(qtx):
    qtx.set_bc1_encoder(LANE)
    yield qtx
    qtx.set_bc1_encoder(0)


def _params(q, build, color_mode, interp):
    """build 'light': level 10's preset (the level 0-12 kernel); 'any': level 9
    with one endpoint-search round (the any-settings kernel, still below the
    queued kernel's cost)."""
    if build == 'light':
        return q.params(10, color_mode, interp)
    p = q.params(9, color_mode, interp)
    p.search_rounds = 1
    return p


def _solid_heavy(w, h, seed, frac=0.5):
    """The block mixture with a fraction of its blocks made solid (one RGBA value)."""
    img = synth.mixture_rgba(w, h, seed=seed, near_black=0.1)
    rng = np.random.default_rng(seed)
    bw, bh = (w + 3) // 4, (h + 3) // 4
    pick = rng.random((bh, bw)) < frac
    for by, bx in zip(*np.nonzero(pick)):
        img[4 * by:4 * by + 4, 4 * bx:4 * bx + 4] = rng.integers(0, 256, 4, dtype=np.uint8)
    return img


def _encode_device(q, hip, img, p, stream, bc3=False):
    h, w = img.shape[:2]
    nb = ((w + 3) // 4) * ((h + 3) // 4)
    d_in = hip.Buffer.from_array(img)
    d_out = hip.Buffer(nb * (16 if bc3 else 8))
    fn = q.bc3_encode_device if bc3 else q.bc1_encode_device
    fn(p, d_in.ptr, w, h, d_out.ptr, stream.handle)
    return d_in, d_out


@pytest.mark.parametrize('build', ['light', 'any'])
@pytest.mark.parametrize('interp', [0, 1, 2, 3])
@pytest.mark.parametrize('bc3', [False, True])
def test_builds(lane, checker, build, interp, bc3):
    img = _solid_heavy(196, 88, seed=100 + interp + 4 * bc3)
    if bc3:
        img[..., 3] = synth.alpha_channel(196, 88, seed=110 + interp)
        img[:16, :16, 3] = 77  # solid blocks whose alpha is not 255
    else:
        img[..., 3] = 255
    p = _params(lane, build, 4 if bc3 else 0x13, interp)
    if bc3:
        got, want = lane.bc3_encode(img, p), checker.bc3_encode(img, p)
    else:
        got, want = lane.bc1_encode(img, p), checker.bc1_encode(img, p)
    assert (got == want).all()


@pytest.mark.parametrize('level', [0, 6, 7, 8, 9, 10, 11, 12])
def test_levels(lane, checker, level):
    img = _solid_heavy(128, 64, seed=120 + level, frac=0.3)
    p = lane.params(level, 0x13, level % 4)
    assert (lane.bc1_encode(img, p) == checker.bc1_encode(img, p)).all()


def test_relaunch_same_stream(lane, checker):
    from quicktex_amd import hip
    q = lane
    big = _solid_heavy(512, 256, seed=131)
    small = _solid_heavy(96, 40, seed=132)
    p = q.params(10, 0x13, 0)
    want_big = checker.bc1_encode(big, p)
    want_small = checker.bc1_encode(small, p)
    s = hip.Stream()
    for img, want in ((big, want_big), (big, want_big), (small, want_small), (big, want_big)):
        d_in, d_out = _encode_device(q, hip, img, p, s)
        s.synchronize()
        got = d_out.download().reshape(-1)
        assert (got == want).all(), img.shape
    s.destroy()


def test_many_streams(lane, checker):
    """More launches in flight than queue sets, on as many streams."""
    from quicktex_amd import hip
    q = lane
    img = _solid_heavy(256, 128, seed=133)
    p = q.params(8, 4, 2)
    want = checker.bc1_encode(img, p)
    streams = [hip.Stream() for _ in range(72)]
    held = [_encode_device(q, hip, img, p, s) for s in streams]
    hip.synchronize()
    for i, (d_in, d_out) in enumerate(held):
        assert (d_out.download().reshape(-1) == want).all(), 'stream %d' % i
    for s in streams:
        s.destroy()


def test_many_launches_one_stream(lane, checker):
    """More launches queued on one stream than the context has queue sets: each
    takes a set last used on the same stream, whose words stream order has
    returned to zero."""
    from quicktex_amd import hip
    q = lane
    img = _solid_heavy(256, 64, seed=134)
    p = q.params(9, 0x13, 3)
    want = checker.bc1_encode(img, p)
    nb = (256 // 4) * (64 // 4)
    s = hip.Stream()
    d_in = hip.Buffer.from_array(img)
    n = 100
    d_out = hip.Buffer(n * nb * 8)
    d_out.zero(0xEE)
    for k in range(n):
        q.bc1_encode_device(p, d_in.ptr, 256, 64, d_out.ptr + k * nb * 8, s.handle)
    s.synchronize()
    got = d_out.download().reshape(n, nb * 8)
    bad = [k for k in range(n) if not (got[k] == want).all()]
    assert not bad, 'launches %s differ' % bad[:8]
    s.destroy()


def test_automatic_route(qtx, checker):
    """1024x512 (32768 blocks) at level 8: the automatic choice takes the static
    lane kernel, and with it the list."""
    img = _solid_heavy(1024, 512, seed=135, frac=0.25)
    p = qtx.params(8, 0x13, 1)
    got = qtx.bc1_encode(img, p)
    rows = 64
    want = checker.bc1_encode(np.ascontiguousarray(img[:rows]), p)
    assert (got[:want.size] == want).all()
    qtx.set_bc1_encoder(2)  # the 16-lane encoder, for every block
    try:
        assert (got == qtx.bc1_encode(img, p)).all()
    finally:
        qtx.set_bc1_encoder(0)


@pytest.mark.parametrize('kind', ['all_solid', 'none_solid', 'one_block', 'one_solid_block', 'ragged'])
@pytest.mark.parametrize('bc3', [False, True])
def test_list_edges(lane, checker, kind, bc3):
    rng = np.random.default_rng(['all_solid', 'none_solid', 'one_block', 'one_solid_block', 'ragged'].index(kind) * 2 + bc3)
    if kind == 'all_solid':
        img = np.repeat(np.repeat(rng.integers(0, 256, (32, 48, 4), dtype=np.uint8), 4, 0), 4, 1)
    elif kind == 'none_solid':
        img = rng.integers(0, 256, (128, 192, 4), dtype=np.uint8)
    elif kind == 'one_block':
        img = rng.integers(0, 256, (4, 4, 4), dtype=np.uint8)
    elif kind == 'one_solid_block':
        img = np.broadcast_to(rng.integers(0, 256, 4, dtype=np.uint8), (3, 2, 4)).copy()
    else:
        img = _solid_heavy(203, 77, seed=136)
    if not bc3:
        img[..., 3] = 255
    p = lane.params(10, 4, 0)
    if bc3:
        got, want = lane.bc3_encode(img, p), checker.bc3_encode(img, p)
    else:
        got, want = lane.bc1_encode(img, p), checker.bc1_encode(img, p)
    assert (got == want).all()
