"""GPU: the BC1/BC3 lane encoder that deals blocks from work queues.

Each launch of the queued kernel borrows one of the context's 64 sets of queue
words until it completes, and leaves the words zero.  Checked here, bit-exact
against the CPU checker: relaunches on one stream (the words must be back at
zero, also when a smaller texture follows a larger one), many launches in
flight on many streams at once (each holds its own set; past 64 the queue-less
kernel takes over), and the dealing's edge cases: all blocks solid, none solid,
one block, ragged sizes.  Endpoint searches that run past the park stop are
parked by the queued kernel and finished by the drain kernel: checked at level
18 for every interpolator, BC1 and BC3, relaunched on one stream (the park words
return to zero) and in flight on several streams (one park buffer per set), and
at level 19 on a texture large enough for the automatic choice to take the queued
kernel (against the 16-lane encoder's bytes).
"""
import numpy as np
import pytest

from quicktex_amd import synth

pytestmark = pytest.mark.gpu

LANE = 1  # QTX_BC1_LANE: the lane encoder at every size
LEVEL = 13  # costly enough for the queued kernel (estimated ops per block >= 14000)


@pytest.fixture(scope='module')
def checker():
    from tests._cpu_libs import oracle, reference
    return reference() or oracle()


@pytest.fixture
def lane(qtx):
    qtx.set_bc1_encoder(LANE)
    yield qtx
    qtx.set_bc1_encoder(0)


def _encode_device(q, hip, img, p, stream, bc3=False):
    h, w = img.shape[:2]
    nb = ((w + 3) // 4) * ((h + 3) // 4)
    d_in = hip.Buffer.from_array(img)
    d_out = hip.Buffer(nb * (16 if bc3 else 8))
    fn = q.bc3_encode_device if bc3 else q.bc1_encode_device
    fn(p, d_in.ptr, w, h, d_out.ptr, stream.handle)
    return d_in, d_out


def test_relaunch_same_stream(lane, checker):
    from quicktex_amd import hip
    q = lane
    big = synth.mixture_rgba(512, 256, seed=21, near_black=0.1)
    small = synth.mixture_rgba(96, 40, seed=22, near_black=0.1)
    p = q.params(LEVEL, 0x13, 0)
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
    img = synth.mixture_rgba(256, 128, seed=23, near_black=0.1)
    p = q.params(LEVEL, 4, 2)
    want = checker.bc1_encode(img, p)
    streams = [hip.Stream() for _ in range(72)]
    held = [_encode_device(q, hip, img, p, s) for s in streams]
    hip.synchronize()
    for i, (d_in, d_out) in enumerate(held):
        assert (d_out.download().reshape(-1) == want).all(), 'stream %d' % i
    for s in streams:
        s.destroy()


def test_concurrent_large(lane, checker):
    """Two large launches on two streams at once, then again on swapped inputs."""
    from quicktex_amd import hip
    q = lane
    a = synth.mixture_rgba(1024, 512, seed=24, near_black=0.1)
    b = synth.mixture_rgba(768, 512, seed=25, near_black=0.2)
    p = q.params(LEVEL, 0x13, 1)
    rows = 64  # the checker's share: the first 64 rows of each
    want_a = checker.bc1_encode(np.ascontiguousarray(a[:rows]), p)
    want_b = checker.bc1_encode(np.ascontiguousarray(b[:rows]), p)
    s1, s2 = hip.Stream(), hip.Stream()
    for x, y in ((a, b), (b, a)):
        hx = _encode_device(q, hip, x, p, s1)
        hy = _encode_device(q, hip, y, p, s2)
        s1.synchronize()
        s2.synchronize()
        for img, (_, d_out) in ((x, hx), (y, hy)):
            got = d_out.download().reshape(-1)
            want = want_a if img is a else want_b
            assert (got[:want.size] == want).all(), img.shape
            # every block written: the same texture through the host entry point
            assert (got == q.bc1_encode(img, p)).all(), img.shape
    s1.destroy()
    s2.destroy()


@pytest.mark.parametrize('kind', ['all_solid', 'none_solid', 'one_block', 'ragged'])
@pytest.mark.parametrize('bc3', [False, True])
def test_dealing_edges(lane, checker, kind, bc3):
    rng = np.random.default_rng(['all_solid', 'none_solid', 'one_block', 'ragged'].index(kind) * 2 + bc3)
    if kind == 'all_solid':
        img = np.repeat(np.repeat(rng.integers(0, 256, (32, 48, 4), dtype=np.uint8), 4, 0), 4, 1)
    elif kind == 'none_solid':
        img = rng.integers(0, 256, (128, 192, 4), dtype=np.uint8)
    elif kind == 'one_block':
        img = rng.integers(0, 256, (4, 4, 4), dtype=np.uint8)
    else:
        img = synth.mixture_rgba(203, 77, seed=26, near_black=0.1)
    if not bc3:
        img[..., 3] = 255
    p = lane.params(LEVEL, 4, 0)
    if bc3:
        got, want = lane.bc3_encode(img, p), checker.bc3_encode(img, p)
    else:
        got, want = lane.bc1_encode(img, p), checker.bc1_encode(img, p)
    assert (got == want).all()


# Endpoint searches longer than the park stop (round 34) are parked by the queued
# kernel and finished by the drain kernel; level 18 searches up to 256 rounds.
PARK_LEVEL = 18


@pytest.mark.parametrize('interp', [0, 1, 2, 3])
@pytest.mark.parametrize('bc3', [False, True])
def test_parked_searches(lane, checker, interp, bc3):
    img = synth.mixture_rgba(160, 96, seed=40 + interp, near_black=0.1)
    if bc3:
        img[..., 3] = synth.alpha_channel(160, 96, seed=50 + interp)
    p = lane.params(PARK_LEVEL, 4 if bc3 else 0x13, interp)
    if bc3:
        got, want = lane.bc3_encode(img, p), checker.bc3_encode(img, p)
    else:
        got, want = lane.bc1_encode(img, p), checker.bc1_encode(img, p)
    assert (got == want).all()


def test_parked_relaunch_and_streams(lane, checker):
    """The park words return to zero after each launch (a smaller texture after a
    larger one on one stream), and launches in flight on several streams each park
    into their own set's buffer."""
    from quicktex_amd import hip
    q = lane
    big = synth.mixture_rgba(256, 192, seed=61, near_black=0.1)
    small = synth.mixture_rgba(72, 36, seed=62, near_black=0.1)
    p = q.params(PARK_LEVEL, 0x13, 2)
    want_big, want_small = checker.bc1_encode(big, p), checker.bc1_encode(small, p)
    s = hip.Stream()
    for img, want in ((big, want_big), (small, want_small), (big, want_big)):
        d_in, d_out = _encode_device(q, hip, img, p, s)
        s.synchronize()
        assert (d_out.download().reshape(-1) == want).all(), img.shape
    s.destroy()
    streams = [hip.Stream() for _ in range(6)]
    held = [_encode_device(q, hip, big if i % 2 == 0 else small, p, st) for i, st in enumerate(streams)]
    hip.synchronize()
    for i, (d_in, d_out) in enumerate(held):
        want = want_big if i % 2 == 0 else want_small
        assert (d_out.download().reshape(-1) == want).all(), 'stream %d' % i
    for st in streams:
        st.destroy()


@pytest.mark.parametrize('interp', [0, 2])
def test_level19_large_texture_both_encoders(qtx, interp):
    """Level 19 (exhaustive cluster fit) on a 2048x2048 texture (262144 blocks):
    routed automatically to the queued lane kernel with parking and the drain;
    its bytes must equal the 16-lane encoder's (which the parity tests check
    against the CPU at every level on smaller textures)."""
    from quicktex_amd import hip
    img = synth.mixture_rgba(2048, 2048, seed=70 + interp, near_black=0.1)
    p = qtx.params(19, 0x13, interp)
    s = hip.Stream()
    outs = {}
    for enc in (0, 2):  # automatic (lane at this size), 16-lane
        qtx.set_bc1_encoder(enc)
        d_in, d_out = _encode_device(qtx, hip, img, p, s)
        s.synchronize()
        outs[enc] = d_out.download()
    qtx.set_bc1_encoder(0)
    s.destroy()
    assert (outs[0] == outs[2]).all()


def test_many_launches_one_stream_reuse_sets(lane, checker):
    """More launches queued on one stream than the context has queue sets: later
    launches take sets whose last launch was on the same stream (stream order
    frees them in time) and keep the queued kernel and its parked searches."""
    from quicktex_amd import hip
    q = lane
    img = synth.mixture_rgba(256, 64, seed=29, near_black=0.1)
    p = q.params(18, 0x13, 2)
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
