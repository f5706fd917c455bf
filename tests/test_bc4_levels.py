"""The BC4 encoder's selector level in float (bc4_encode_values in
quicktex_amd/csrc/kernels.hip) against the reference's count of thresholds
(quicktex/s3tc/bc4/BC4Encoder.cpp:30-71), for every block range and value.

The kernel computes floor(fma(x, r, 2^-11)) + 1, capped at 7, with
x = 14 (v - min) + 4 - delta and r the correctly rounded 1 / (2 delta); here the
fma is emulated by one float64 evaluation rounded once to float32 (x r is exact
in float64)."""
import numpy as np


def test_level_by_quotient_matches_threshold_count():
    eps = np.float32(2.0 ** -11)
    for delta in range(1, 256):
        r = np.float32(1.0) / np.float32(2 * delta)
        d = np.arange(delta + 1)
        x = 14 * d + 4 - delta
        # the reference: value = 14 v, bias = 4 - 14 min, level = #{c < 7 : value >= delta (1 + 2c) - bias}
        value = 14 * (d + 0)  # v - min = d; min cancels
        count = np.zeros_like(d)
        for c in range(7):
            count += (value >= delta * (1 + 2 * c) - 4) .astype(count.dtype)
        t = (x.astype(np.float32).astype(np.float64) * np.float64(r) + np.float64(eps)).astype(np.float32)
        level = np.minimum(np.floor(t).astype(np.int64) + 1, 7)
        assert (level == count).all(), delta


def test_level_quotient_margin():
    """A plausible slip -- no 2^-11 term -- fails on exact quotients (x = 2 delta k)."""
    bad = 0
    for delta in range(1, 256):
        r = np.float32(1.0) / np.float32(2 * delta)
        x = np.arange(0, 14 * delta + 4 - delta + 1, 2 * delta)
        t = (x.astype(np.float64) * np.float64(r)).astype(np.float32)
        bad += int((np.floor(t).astype(np.int64) != x // (2 * delta)).sum())
    assert bad > 0
