"""The block-row division of the HBM-bound kernels (RowDiv in
quicktex_amd/csrc/kernels_dev.h): floor(gid / bw) as mulhi(gid, m) >> s with
l = ceil(log2 bw), m = floor(2^(31+l) / bw) + 1, s = l - 1 (Granlund and
Montgomery 1994, theorem 4.2, N = 31).  Checked against integer division over
every width up to 4096, powers of two and their neighbours up to 2^27 (the
largest block width a texture of at most 2^31 texels has), random widths, and
for each the dividends where a wrong multiplier fails first: the multiples of
bw and their predecessors, near 0 and near 2^31.
"""
import numpy as np


def row_div(bw):
    if bw <= 1:
        return 0, 0
    l = (bw - 1).bit_length()  # ceil(log2 bw)
    return (1 << (31 + l)) // bw + 1, l - 1


def rows(gid, bw):
    m, s = row_div(bw)
    if m == 0:
        return gid
    assert m < 1 << 32
    return ((gid * m) >> 32) >> s


def _dividends(bw, rng):
    top = (1 << 31) - 1
    k_hi = top // bw
    ks = np.unique(np.concatenate([np.arange(0, 40), np.arange(max(0, k_hi - 40), k_hi + 1),
                                   rng.integers(0, k_hi + 1, 200)]))
    out = set()
    for k in ks.tolist():
        for n in (k * bw - 1, k * bw, k * bw + 1, k * bw + bw - 1):
            if 0 <= n <= top:
                out.add(n)
    out.update((0, 1, top, top - 1))
    return np.array(sorted(out), dtype=object)


def _widths():
    rng = np.random.default_rng(7)
    ws = set(range(1, 4097))
    for e in range(1, 28):
        ws.update(((1 << e) - 1, 1 << e, (1 << e) + 1))
    ws.update(int(x) for x in rng.integers(4097, 1 << 27, 300))
    return sorted(ws)


def test_row_div_exact():
    rng = np.random.default_rng(11)
    bad = []
    for bw in _widths():
        for n in _dividends(bw, rng):
            n = int(n)
            if rows(n, bw) != n // bw:
                bad.append((bw, n))
                break
    assert not bad, bad[:5]


def test_multiplier_fits_32_bits():
    for bw in _widths():
        m, s = row_div(bw)
        assert 0 <= m < 1 << 32 and 0 <= s < 32, bw
