"""CPU: the integer cuts that give the selector searches' candidate index in the
lane and 16-lane encoders (quicktex_amd/csrc/bc1_device.h, check2_cuts and
faster_cuts).

The reference's Check2 selector search (quicktex/s3tc/bc1/BC1Encoder.cpp:721-749)
takes, per texel, k = clamp(int(x * f + 0.5f), 1, 3) for the integer projection
x = (p - c0) . axis and f = 4 / (|axis|^2 + 1.25e-6), all in float.  The encoders
instead compare x with two cuts per trial, k = 1 + (x >= t2) + (x >= t3), where
t2 = c2 or c2 + 1 for c2 = ceil(3A / 8) (A = |axis|^2), whichever the rounded
expression at c2 decides, and t3 likewise from ceil(5A / 8).  Here the rule is
checked for every A an axis of 8-bit colours can have, against the float
expression itself at and around each cut, and over random and extreme
projections; a cut one off (or the rule without its correction) fails.

The Faster (and None) search (BC1Encoder.cpp:695-720) takes s = 3 - [2x <= t0] -
[2x < t1] - [2x < t2].  The encoders compare x with the sorted cuts t0 >> 1,
(t1 - 1) >> 1, (t2 - 1) >> 1 and pick the palette entry from three nested
compares; checked against the reference's form over random, tied, unsorted and
extreme thresholds.
"""
import numpy as np

F32 = np.float32
AMAX = 3 * 255 * 255


def _f(A):
    return (F32(4.0) / (A.astype(F32) + F32(1.25e-6))).astype(F32)


def _y(x, f):
    return (x.astype(F32) * f).astype(F32) + F32(0.5)


def _k_ref(x, f):
    return np.clip(np.trunc(_y(x, f)).astype(np.int64), 1, 3)


def _cuts(A, f):
    c2, c3 = (3 * A + 7) >> 3, (5 * A + 7) >> 3
    t2 = np.where(_y(c2, f) >= F32(2.0), c2, c2 + 1)
    t3 = np.where(_y(c3, f) >= F32(3.0), c3, c3 + 1)
    return t2, t3


def test_cuts_are_the_least_projections_reaching_each_index():
    A = np.arange(0, AMAX + 1, dtype=np.int64)
    f = _f(A)
    t2, t3 = _cuts(A, f)
    for t, level in ((t2, 2.0), (t3, 3.0)):
        assert (_y(t, f) >= F32(level)).all()
        assert not (_y(t - 1, f) >= F32(level)).any()
    # around the cuts the index equals the float expression's
    for t in (t2, t3):
        for d in range(-3, 4):
            x = t + d
            assert np.array_equal(1 + (x >= t2) + (x >= t3), _k_ref(x, f))


def test_random_and_extreme_projections():
    rng = np.random.default_rng(0xC2C7)
    A = rng.integers(0, AMAX + 1, 400000)
    A[:8] = [0, 1, 2, 3, 8, 16, AMAX - 1, AMAX]
    f = _f(A)
    x = rng.integers(-AMAX, AMAX + 1, A.size)
    x[:4] = [-AMAX, AMAX, 0, 1]
    t2, t3 = _cuts(A, f)
    assert np.array_equal(1 + (x >= t2) + (x >= t3), _k_ref(x, f))


def test_correction_is_needed():
    """Without the evaluation at c2 / c3 the cuts are off for some A (multiples of 8)."""
    A = np.arange(0, AMAX + 1, dtype=np.int64)
    f = _f(A)
    t2, t3 = _cuts(A, f)
    c2, c3 = (3 * A + 7) >> 3, (5 * A + 7) >> 3
    assert (t2 != c2).sum() > 100 and (t3 != c3).sum() > 100
    assert ((A[t2 != c2] % 8) == 0).all()


def _faster_ref(x, t0, t1, t2):
    d = 2 * x
    return 3 - ((d <= t0).astype(np.int64) + (d < t1) + (d < t2))


def _faster_cuts(x, t0, t1, t2):
    e = np.stack([t0 >> 1, (t1 - 1) >> 1, (t2 - 1) >> 1])
    f0, f1, f2 = np.sort(e, axis=0)
    return np.where(x <= f0, 0, np.where(x <= f1, 1, np.where(x <= f2, 2, 3)))


def test_faster_cuts():
    rng = np.random.default_rng(0xFA57)
    n = 400000
    lim = 2 * AMAX
    t = rng.integers(-lim, lim + 1, (3, n))
    x = rng.integers(-AMAX, AMAX + 1, n)
    # ties between thresholds, thresholds out of order, x at and next to each threshold
    t[1, :n // 8] = t[0, :n // 8]
    t[2, n // 8:n // 4] = t[1, n // 8:n // 4]
    j = rng.integers(0, 3, n)
    near = (t[j, np.arange(n)] >> 1) + rng.integers(-1, 2, n)
    x[n // 4:] = near[n // 4:]
    assert np.array_equal(_faster_cuts(x, *t), _faster_ref(x, *t))
    # off by one in a cut's rounding changes some selectors
    e_bad = (t[1] >> 1)
    bad = np.where(x <= np.minimum(t[0] >> 1, np.minimum(e_bad, (t[2] - 1) >> 1)), 0, 1)
    assert not np.array_equal(bad == 0, _faster_ref(x, *t) == 0)
