"""Test images beyond the synthetic configs of quicktex_amd.synth."""
import numpy as np


def stress_rgba(w, h, seed):
    """Blocks of 2-3 distinct colours: greyscale, near-black, random."""
    rng = np.random.default_rng(seed)
    img = np.zeros((h, w, 4), np.uint8)
    img[..., 3] = 255
    for by in range(h // 4):
        for bx in range(w // 4):
            kind = rng.integers(0, 5)
            cols = rng.integers(0, 256, (3, 3))
            if kind == 0:
                cols[:, :] = cols[:, :1]
            if kind == 1:
                cols[1] = rng.integers(0, 8, 3)
            pick = rng.integers(0, 3 if kind > 2 else 2, 16)
            img[by * 4:by * 4 + 4, bx * 4:bx * 4 + 4, :3] = cols[pick].reshape(4, 4, 3)
    return img
