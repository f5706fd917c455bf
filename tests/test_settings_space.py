"""The whole BC1 settings space on the CPU: oracle against the compiled reference.

The presets (every level x colour mode x interpolator) are pinned elsewhere;
here the settings are drawn from the space the reference's advanced properties
span (quicktex/s3tc/bc1/_bindings.cpp:134-164) -- ErrorMode.None and every
error mode x endpoint mode pair, the two-pass flags on and off, exhaustive
cluster fit, search rounds at the loop's exit boundaries, orderings and power
iterations at and past their clamps -- and run through the oracle, the compiled
reference (oracle/_ref) and, where the least-squares endpoint mode runs, the
zero-initialising reference build (the g++ build reads an uninitialised
accumulator there: BC1Encoder.cpp:510-515).

The same draws are in tests/golden/settings_vectors.npz, written by the
reference, so the oracle stays pinned where oracle/_ref is not built.
"""
import ctypes
import os

import numpy as np
import pytest

from quicktex_amd import synth
from tests import _settings
from tests._images import stress_rgba

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden', 'settings_vectors.npz')
SEED = 0x5E77


def image():
    mix = synth.mixture_rgba(64, 16, seed=0x5E7700, near_black=0.2)
    return np.concatenate([mix, stress_rgba(64, 16, 0x5E7701)], axis=0)  # 64 x 32: 128 blocks


def draws():
    return _settings.draws(SEED, 12)


@pytest.fixture(scope='module')
def lib():
    from tests._qtx import Qtx
    return Qtx().lib


def test_product_clamps_equal_reference_setters(lib, reference):
    """qtx_bc1_params_clamp leaves what the reference's setters leave."""
    for name, raw in draws():
        mine = _settings.copy(raw)
        assert lib.qtx_bc1_params_clamp(ctypes.byref(mine)) == 0
        assert mine.as_tuple() == reference.apply(raw).as_tuple(), (name, _settings.describe(raw))


def test_oracle_equals_reference_over_settings(oracle, reference, reference_zi):
    img = image()
    for name, raw in draws():
        p = reference.apply(raw)
        ref = reference_zi if p.endpoint_mode == 0 else reference
        got, want = oracle.bc1_encode(img, p), ref.bc1_encode(img, p)
        bad = np.nonzero((got.reshape(-1, 8) != want.reshape(-1, 8)).any(axis=1))[0]
        assert len(bad) == 0, '%s (%s): blocks %s' % (name, _settings.describe(p), bad[:6].tolist())


def test_oracle_equals_reference_bc3_over_settings(oracle, reference, reference_zi):
    img = image()
    rng = np.random.default_rng(SEED + 1)
    img[..., 3] = rng.integers(0, 256, img.shape[:2], dtype=np.uint8)
    img[:8, :16, 3] = 255
    for name, raw in draws()[::3]:
        raw.color_mode = 4  # BC3's colour encoder is FourColor (BC3Encoder.h:37-38)
        p = reference.apply(raw)
        ref = reference_zi if p.endpoint_mode == 0 else reference
        assert (oracle.bc3_encode(img, p) == ref.bc3_encode(img, p)).all(), (name, _settings.describe(p))


def test_zero_init_build_equals_gcc_build_where_defined(reference, reference_zi):
    """Off the least-squares endpoint mode the two reference builds agree, so the
    zero-initialising build is the reference wherever the reference is defined."""
    img = image()
    for name, raw in draws()[::2]:
        p = reference.apply(raw)
        if p.endpoint_mode == 0:
            continue
        assert (reference.bc1_encode(img, p) == reference_zi.bc1_encode(img, p)).all(), name


@pytest.mark.parametrize('interp', [0, 1, 2, 3])
def test_level1_every_interpolator(oracle, reference_zi, interp):
    """Level 1 (EndpointMode.LeastSquares) for every interpolator and colour mode."""
    img = synth.mixture_rgba(256, 64, seed=0x1E1 + interp, near_black=0.2)
    img = np.concatenate([img, stress_rgba(256, 32, 0x1E5 + interp)], axis=0)
    for mode in (4, 3, 0x13):
        p = oracle.params(1, mode, interp)
        assert (oracle.bc1_encode(img, p) == reference_zi.bc1_encode(img, p)).all(), mode


def test_oracle_equals_settings_vectors(oracle):
    """The reference's own outputs for the same draws (no compiled reference needed)."""
    v = np.load(GOLDEN)
    img, params, blocks = v['input'], v['params'], v['blocks']
    assert (img == image()).all()
    from quicktex_amd.capi import BC1Params
    for k in range(params.shape[0]):
        p = BC1Params(*[int(x) for x in params[k]])
        assert (oracle.bc1_encode(img, p) == blocks[k]).all(), (k, _settings.describe(p))
