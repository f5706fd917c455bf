"""GPU: the whole BC1 settings space through both encoders and the automatic choice.

The draws of tests/_settings.py (ErrorMode.None, every error mode x endpoint
mode, two-pass flags, exhaustive, search rounds at the loop's exit boundaries,
orderings and power iterations at and past their clamps) run through the C ABI
with the one-block-per-lane encoder forced (its static kernel below 10k
estimated ops per block, else the queued kernel in its any-settings build, or
its level 13-18 build for draws that share those settings, with parked
searches past round 34, or its level 0-12 build), the 16-lane encoder forced,
and the automatic choice,
each bit-exact against the compiled reference (the zero-initialising build
where the least-squares endpoint mode runs).  A few draws also run on a
2048 x 1024 texture, where the automatic choice takes the queued kernel:
every block against the 16-lane encoder's bytes, and every 16th block row, the
first and the last against the reference.
"""
import numpy as np
import pytest

from quicktex_amd import synth
from tests import _settings
from tests._images import stress_rgba

pytestmark = pytest.mark.gpu

SEED = 0x6E77
ENCODERS = {'auto': 0, 'lane': 1, 'group': 2}


def _draws():
    return _settings.draws(SEED, 6)


@pytest.fixture(scope='module')
def image():
    mix = synth.mixture_rgba(160, 64, seed=0x6E7700, near_black=0.2)
    return np.concatenate([mix, stress_rgba(160, 32, 0x6E7701)], axis=0)  # 160 x 96: 960 blocks


def _clamped(qtx, raw):
    import ctypes
    p = _settings.copy(raw)
    qtx.check(qtx.lib.qtx_bc1_params_clamp(ctypes.byref(p)))
    return p


def _checker(p, oracle):
    from tests._cpu_libs import checker_for
    return checker_for(p) or oracle


@pytest.mark.parametrize('which', sorted(ENCODERS))
@pytest.mark.parametrize('k', range(len(_settings.draws(SEED, 6))))
def test_settings_draw(qtx, oracle, image, k, which):
    name, raw = _draws()[k]
    p = _clamped(qtx, raw)
    want = _checker(p, oracle).bc1_encode(image, p)
    qtx.set_bc1_encoder(ENCODERS[which])
    try:
        got = qtx.bc1_encode(image, p)
    finally:
        qtx.set_bc1_encoder(0)
    bad = np.nonzero((got.reshape(-1, 8) != want.reshape(-1, 8)).any(axis=1))[0]
    assert len(bad) == 0, '%s %s (%s): %d blocks, first %s' % (name, which, _settings.describe(p), len(bad),
                                                               bad[:4].tolist())


def _big_draws():
    """Draws whose automatic choice at 131072 blocks is the queued kernel
    (estimated ops per block of at least 10k and at most 2 x 65536): two with the
    level 13-18 settings and two without, with searches long enough to park."""
    out = []
    rng = np.random.default_rng(SEED + 7)
    for hq in (True, True, False, False):
        p = _settings._draw(rng, 'hq' if hq else 'free', False)
        p.search_rounds = int(rng.choice([35, 100, 256]))
        p.orderings4 = int(rng.integers(16, 64))
        p.orderings3 = int(rng.integers(1, 16))
        p.two_cf_passes = 0
        if not hq:
            p.error_mode = int(rng.choice([1, 2]))
        out.append(p)
    return out


@pytest.mark.parametrize('k', range(4))
def test_settings_auto_queued_texture(qtx, oracle, k):
    p = _clamped(qtx, _big_draws()[k])
    w, h = 2048, 1024
    img = synth.mixture_rgba(w, h, seed=0x6E7710 + k, near_black=0.1)
    print('settings %s at %dx%d' % (_settings.describe(p), w, h), flush=True)
    auto = qtx.bc1_encode(img, p).reshape(h // 4, w // 4, 8)
    qtx.set_bc1_encoder(2)
    try:
        group = qtx.bc1_encode(img, p).reshape(h // 4, w // 4, 8)
    finally:
        qtx.set_bc1_encoder(0)
    assert (auto == group).all(), 'automatic choice differs from the 16-lane encoder'
    rows = sorted(set(list(range(0, h // 4, 16)) + [h // 4 - 1]))
    ref = _checker(p, oracle)
    for r in rows:
        want = ref.bc1_encode(img[4 * r:4 * r + 4], p).reshape(w // 4, 8)
        assert (auto[r] == want).all(), 'block row %d' % r
