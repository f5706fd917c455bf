"""The BC4 encoder's selector (bc4_encode_values in quicktex_amd/csrc/kernels.hip)
against the reference's count of thresholds (quicktex/s3tc/bc4/BC4Encoder.cpp:30-71),
for every block range and value.

The kernel takes floor(y3 / d) = level + 1 (capped at 8..10 for level 7) as the
high word of the 24 x 24-bit product (y3 2^10) m, y3 = 14 (v - min) + 4 + 3 delta,
d = 2 delta and m = ceil(2^22 / d) (= (2^22 - 1) // d + 1), and the selector
from it with a byte permute (v_perm_b32) of a table whose index 0 is zero and
whose indices 8-10 give 0.
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


def _perm(src0, src1, sel):
    """v_perm_b32: result byte j from selector byte j -- 0-3 bytes of src1, 4-7 bytes
    of src0, 8 / 9 / 10 / 11 the sign (0x00 or 0xFF) of src1 byte 1 / 3, src0 byte 1 /
    3, 12 zero, 13 and above 0xFF."""
    table = [(src1 >> (8 * j)) & 0xFF for j in range(4)] + [(src0 >> (8 * j)) & 0xFF for j in range(4)]
    out = 0
    for j in range(4):
        s = (sel >> (8 * j)) & 0xFF
        if s < 8:
            b = table[s]
        elif s < 12:
            b = 0xFF if table[{8: 1, 9: 3, 10: 5, 11: 7}[s]] & 0x80 else 0x00
        else:
            b = 0x00 if s == 12 else 0xFF
        out |= b << (8 * j)
    return out


def test_selector_by_reciprocal_matches_threshold_count():
    order = [1, 7, 6, 5, 4, 3, 2, 0]  # BC4Encoder.cpp:44
    qs = set()
    for delta in range(1, 256):
        d = 2 * delta
        m = ((1 << 22) - 1) // d + 1
        assert m <= (1 << 21)
        for mn in sorted({0, 1, (255 - delta) // 2, 255 - delta}):
            k = ((4 + 3 * delta - 14 * mn) << 10) & 0xFFFFFFFF
            for v in range(mn, mn + delta + 1):
                y = (((14 << 10) * v) + k) & 0xFFFFFFFF  # __umul24(v, 14 << 10) + k
                assert y < (1 << 24)                      # a 24-bit operand, as v_mul_hi_u32_u24 reads it
                q = (y * m) >> 32
                qs.add(q)
                count = sum(14 * v >= delta * (1 + 2 * c) - (4 - 14 * mn) for c in range(7))
                assert _perm(0x02030405, 0x06070100, q) == order[count], (delta, mn, v)
    assert qs == set(range(1, 11))


def test_reciprocal_correction():
    """m from any estimate within one of floor((2^22 - 1) / d) equals that floor + 1."""
    for d in range(2, 511, 2):
        exact = ((1 << 22) - 1) // d
        for f0 in (exact - 1, exact, exact + 1):
            f, r = f0, ((1 << 22) - 1) - f0 * d
            if r < 0:
                f, r = f - 1, r + d
            if r >= d:
                f += 1
            assert f == exact, (d, f0)


def test_selector_reciprocal_rounded_down_fails():
    """A plausible slip -- m rounded down, floor(2^22 / (2 delta)) -- misses exact quotients."""
    bad = 0
    for delta in range(1, 256):
        m = (1 << 22) // (2 * delta)
        for d in range(delta + 1):
            y = 14 * d + 4 + 3 * delta
            bad += ((y * m) >> 22) != y // (2 * delta)
    assert bad > 0
