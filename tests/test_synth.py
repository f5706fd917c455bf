"""The synthetic workloads are deterministic and independent of tiling."""
import hashlib

import numpy as np

from quicktex_amd import synth


def test_mixture_deterministic_and_tiling_independent():
    a = synth.mixture_rgba(512, 64, seed=1, near_black=0.1)
    b = synth.mixture_rgba(512, 64, seed=1, near_black=0.1)
    assert (a == b).all()
    # the top rows of a taller texture are the same blocks
    c = synth.mixture_rgba(512, 128, seed=1, near_black=0.1)
    assert (c[:64] == a).all()
    assert a[..., 3].min() == 255


def test_mixture_block_kinds():
    a = synth.mixture_rgba(1024, 1024, seed=0x5EED0002, near_black=0.10).reshape(256, 4, 256, 4, 4)
    blocks = a.transpose(0, 2, 1, 3, 4).reshape(-1, 16, 4)
    solid = (blocks == blocks[:, :1]).all(axis=(1, 2)).mean()
    dark = (blocks[..., :3].max(axis=2) < 8).any(axis=1).mean()
    assert 0.20 < solid < 0.26
    assert dark > 0.09


def test_fixed_hashes():
    """Pinned bytes: the benchmark input is the same on every machine."""
    h = hashlib.sha256(synth.mixture_rgba(256, 256).tobytes()).hexdigest()
    h2 = hashlib.sha256(synth.smooth_channels(64, 64).tobytes()).hexdigest()
    assert (h[:16], h2[:16]) == (H_MIX, H_SMOOTH)


H_MIX = '7b73a3f5c6fb0b56'
H_SMOOTH = 'b655f944dc12c3dd'
