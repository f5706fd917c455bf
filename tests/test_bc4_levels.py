"""The BC4 encoder's selector (bc4_encode_values in quicktex_amd/csrc/kernels.hip)
against the reference's count of thresholds (quicktex/s3tc/bc4/BC4Encoder.cpp:30-71),
for every block range and value.

The kernel takes the level as the high word of y m, y = 14 (v - min) + 4 + delta
and m = ceil(2^32 / (2 delta)) (= (2^32 - 1) // (2 delta) + 1), and the selector
from the level with a byte permute whose out-of-table indices 8 and 9 give 0.
The first tests pin an earlier float form of the same quotient, kept as a
cross-check of the closed form."""
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


def _perm_lut(q):
    """v_perm_b32(0x00020304, 0x05060701, 0x0C0C0C00 | q) byte 0, for q in 0..11:
    bytes 0-7 of (src1, src0), 8-11 the sign bytes of src1 byte 1 / 3, src0 byte 1 / 3."""
    src1, src0 = 0x05060701, 0x00020304
    table = [(src1 >> (8 * j)) & 0xFF for j in range(4)] + [(src0 >> (8 * j)) & 0xFF for j in range(4)]
    if q < 8:
        return table[q]
    byte = {8: table[1], 9: table[3], 10: table[5], 11: table[7]}[q]
    return 0xFF if byte & 0x80 else 0x00


def test_selector_by_reciprocal_matches_threshold_count():
    order = [1, 7, 6, 5, 4, 3, 2, 0]  # BC4Encoder.cpp:44
    qmax = 0
    for delta in range(1, 256):
        for mn in (0, 1, 255 - delta):
            m = (0xFFFFFFFF // (2 * delta)) + 1
            k = (4 + delta - 14 * mn) & 0xFFFFFFFF
            for v in range(mn, mn + delta + 1):
                y = (14 * v + k) & 0xFFFFFFFF
                q = (y * m) >> 32
                qmax = max(qmax, q)
                count = sum(14 * v >= delta * (1 + 2 * c) - (4 - 14 * mn) for c in range(7))
                assert _perm_lut(q) == order[count], (delta, mn, v)
    assert qmax == 9


def test_selector_reciprocal_rounded_down_fails():
    """A plausible slip -- m rounded down, floor(2^32 / (2 delta)) -- misses exact quotients."""
    bad = 0
    for delta in range(1, 256):
        m = (1 << 32) // (2 * delta)
        for d in range(delta + 1):
            y = 14 * d + 4 + delta
            bad += ((y * m) >> 32) != y // (2 * delta)
    assert bad > 0
