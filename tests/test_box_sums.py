"""The box-sum form of the mip horizontal pass (k_mip_hrows in
quicktex_amd/csrc/resample.hip).  For a shrink by r = 2^a the interior output o
of a row has the window [o r - r/2, o r + 3r/2) and the weights
(2t + 1) << (21 - 2a), rising over t < r and mirrored after (the closed form of
Pillow's normalised triangle weights in 22-bit fixed point, which the GPU tests
pin against the Pillow-checked oracle).  The kernel replaces the 2r-tap sum by
sums over four boxes of h = r/2 texels, S (texel sum) and M (texel times its
offset in the box), built level after level (S_2h = S_lo + S_hi,
M_2h = M_lo + M_hi + h S_hi).  Checked here on random and extreme rows against
the direct weighted sum, exactly, for every shrink the kernel uses (16 .. 1024)
and for 2, 4 and 8, together with the 31-bit bound the kernel's int32
arithmetic relies on."""
import numpy as np


def pyramid(row, top):
    S, M = {1: row.astype(np.int64)}, {1: np.zeros(row.size, np.int64)}
    h = 1
    while h < top:
        s, m = S[h], M[h]
        S[2 * h] = s[0::2] + s[1::2]
        M[2 * h] = m[0::2] + m[1::2] + h * s[1::2]
        h *= 2
    return S, M


def box_form(S, M, h, o):
    q = 2 * o - 1
    s, m = S[h][q:q + 4], M[h][q:q + 4]
    return (2 * m[0] + s[0] + 2 * m[1] + (2 * h + 1) * s[1] + (4 * h - 1) * s[2] - 2 * m[2]
            + (2 * h - 1) * s[3] - 2 * m[3])


def direct(row, r, o):
    t = np.arange(2 * r)
    w = np.where(t < r, 2 * t + 1, 4 * r - 2 * t - 1)
    lo = o * r - r // 2
    return int((w * row[lo:lo + 2 * r].astype(np.int64)).sum())


def _rows():
    rng = np.random.default_rng(5)
    yield rng.integers(0, 256, 4096)
    yield np.full(4096, 255)
    yield np.tile([0, 255], 2048)
    yield (np.arange(4096) * 7) % 256


def test_box_sums_equal_direct_sums():
    for row in _rows():
        S, M = pyramid(row, 512)
        # the width-8 boxes straight from the row, as the kernel sums them
        b = row.reshape(-1, 8).astype(np.int64)
        assert (S[8] == b.sum(1)).all() and (M[8] == (b * np.arange(8)).sum(1)).all()
        for a in range(1, 11):
            r, h = 1 << a, 1 << (a - 1)
            for o in range(1, row.size // r - 1):
                t = box_form(S, M, h, o)
                assert t == direct(row, r, o), (a, o)
                assert 0 <= t and t << (21 - 2 * a) < 1 << 31
