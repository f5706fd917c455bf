"""Deterministic synthetic textures for the benchmark and parity workloads.

Integer-only and seeded per 4x4 block (splitmix64 of seed, texture and block
index), so the same bytes come out on any machine and for any tiling.

Workloads (block mixtures are per 4x4 block):

* ``mixture_rgba(w, h, seed)``: 25% solid colour, 25% two random colours
  (each texel picks one), 50% linear gradient between two uniform random RGB
  endpoints along a random in-block direction plus Gaussian noise (sigma 8,
  Irwin-Hall of four bytes, rounded), clamped to [0, 255]; alpha 255.
  This is the 256x256 BC1 level-5 workload.
* ``mixture_rgba(..., near_black=0.10)``: as above, but 10% of the blocks are
  gradients whose texels are half replaced by near-black texels (every channel
  in [0, 7]), so the three-colour-with-black mode has work.  This is the
  4096x4096 BC1 level-18 workload.
* ``smooth_channels(w, h, seed, n)``: n channels of uniform noise smoothed by
  a 3x3 box filter ``(sum + 4) // 9`` with edge clamp, then 20% of the blocks
  replaced by a per-block constant; placed in channels 0..n-1 of RGBA8 with
  the others (0, ..., 255).  This is the 2048x2048 BC4/BC5 workload.
"""
import os

import numpy as np

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)
_GOLDEN = np.uint64(0x9E3779B97F4A7C15)


def _splitmix(x):
    """splitmix64 finaliser over a uint64 array (wrapping arithmetic)."""
    with np.errstate(over='ignore'):
        z = x + _GOLDEN
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


class _Stream:
    """A counter-based stream of uint64 words per block."""

    def __init__(self, block_seeds):
        self.s = block_seeds
        self.k = 0

    def u64(self):
        with np.errstate(over='ignore'):
            v = _splitmix(self.s + np.uint64(self.k) * _GOLDEN)
        self.k += 1
        return v

    def bytes(self, n):
        """(nblocks, n) uint8 draws."""
        words = [self.u64() for _ in range((n + 7) // 8)]
        b = np.stack(words, axis=1).view(np.uint8).reshape(len(self.s), -1)
        return b[:, :n]


def _block_seeds(seed, texture, nblocks):
    idx = np.arange(nblocks, dtype=np.uint64)
    base = _splitmix(np.array([seed], dtype=np.uint64))[0] ^ _splitmix(np.array([texture + 0x1234], dtype=np.uint64))[0]
    with np.errstate(over='ignore'):
        return _splitmix(idx * np.uint64(0xD1B54A32D192ED03) + base)


def _noise(st, n):
    """(nblocks, n) int32 approx N(0, 8^2): Irwin-Hall of four bytes, rounded."""
    b = st.bytes(4 * n).reshape(-1, n, 4).astype(np.int32)
    s = b.sum(axis=2) - 510  # sigma ~ 147.8
    # scale by 8/147.8 ~= 443/8192, round half away from zero in integers
    num = s * 443
    return np.where(num >= 0, (num + 4096) >> 13, -((-num + 4096) >> 13)).astype(np.int32)


def _blocks_to_image(blocks, w, h):
    bw, bh = (w + 3) // 4, (h + 3) // 4
    img = blocks.reshape(bh, bw, 4, 4, 4).transpose(0, 2, 1, 3, 4).reshape(bh * 4, bw * 4, 4)
    return np.ascontiguousarray(img[:h, :w])


_POOL = None


def _pool():
    """Threads for the generators (numpy releases the GIL inside its kernels)."""
    global _POOL
    if _POOL is None:
        from concurrent.futures import ThreadPoolExecutor
        _POOL = ThreadPoolExecutor(max_workers=min(16, len(os.sched_getaffinity(0))))
    return _POOL


def mixture_rgba(w, h, seed=0x5EED0000, texture=0, near_black=0.0, threads=True):
    """(h, w, 4) uint8 RGBA texture of the block mixture."""
    bw, bh = (w + 3) // 4, (h + 3) // 4
    step = max(1, 16384 // bw)

    def rows(y0):
        n = min(step, bh - y0) * bw
        # seeds depend on the absolute block index, so chunking does not change content
        seeds = _block_seeds(seed, texture, y0 * bw + n)[y0 * bw:]
        return _mixture_blocks(_Stream(seeds), n, near_black)

    starts = list(range(0, bh, step))
    parts = list(_pool().map(rows, starts)) if threads and len(starts) > 1 else [rows(y0) for y0 in starts]
    return _blocks_to_image(np.concatenate(parts), w, h)


def _mixture_blocks(st, nblocks, near_black):
    """(nblocks, 16, 4) uint8 blocks of the mixture."""
    kind_draw = (st.u64() % np.uint64(1000)).astype(np.int32)
    cols = st.bytes(6).astype(np.int32)
    pick = st.bytes(16)
    dirn = (st.u64() % np.uint64(4)).astype(np.int32)
    noise = _noise(st, 48).reshape(nblocks, 16, 3)
    dark = st.bytes(48).reshape(nblocks, 16, 3).astype(np.int32) & 7
    c0 = cols[:, 0:3]
    c1 = cols[:, 3:6]
    out = np.empty((nblocks, 16, 4), dtype=np.uint8)
    out[:, :, 3] = 255
    x = np.tile(np.arange(4), 4)
    y = np.repeat(np.arange(4), 4)
    tpos = np.stack([x * 2, y * 2, x + y, x + (3 - y)], axis=0)
    t = tpos[dirn]
    grad = (c0[:, None, :] * (6 - t[:, :, None]) + c1[:, None, :] * t[:, :, None] + 3) // 6
    grad = np.clip(grad + noise, 0, 255)
    two = np.where((pick[:, :, None] & 1) == 1, c1[:, None, :], c0[:, None, :])
    solid = np.broadcast_to(c0[:, None, :], (nblocks, 16, 3))
    nb = int(round(near_black * 1000))
    rest = 1000 - nb
    t_solid = nb + rest // 4
    t_two = t_solid + rest // 4
    rgb = np.where((kind_draw >= t_two)[:, None, None], grad, 0)
    rgb = np.where(((kind_draw >= t_solid) & (kind_draw < t_two))[:, None, None], two, rgb)
    rgb = np.where(((kind_draw >= nb) & (kind_draw < t_solid))[:, None, None], solid, rgb)
    if nb:
        blk = np.where((pick[:, :, None] & 2) == 2, dark, grad)
        rgb = np.where((kind_draw < nb)[:, None, None], blk, rgb)
    out[:, :, :3] = rgb.astype(np.uint8)
    return out


def smooth_channels(w, h, seed=0x5EED0001, texture=0, nchan=2):
    """(h, w, 4) uint8: nchan smoothed-noise channels, 20% per-block constants."""
    rng_words = _block_seeds(seed, texture, w * h * nchan)
    vals = (rng_words & np.uint64(0xFF)).astype(np.int32).reshape(nchan, h, w)
    pad = np.pad(vals, ((0, 0), (1, 1), (1, 1)), mode='edge')
    acc = np.zeros_like(vals)
    for dy in range(3):
        for dx in range(3):
            acc += pad[:, dy:dy + h, dx:dx + w]
    sm = (acc + 4) // 9
    bw, bh = (w + 3) // 4, (h + 3) // 4
    bseeds = _block_seeds(seed ^ 0xC0C0, texture, bw * bh)
    const_on = ((bseeds % np.uint64(100)) < np.uint64(20)).reshape(bh, bw)
    const_val = ((bseeds >> np.uint64(8)) & np.uint64(0xFFFF)).astype(np.int32).reshape(bh, bw)
    on = np.repeat(np.repeat(const_on, 4, 0), 4, 1)[:h, :w]
    for c in range(nchan):
        cv = (const_val >> (8 * (c % 2))) & 0xFF
        cvi = np.repeat(np.repeat(cv, 4, 0), 4, 1)[:h, :w]
        sm[c] = np.where(on, cvi, sm[c])
    out = np.zeros((h, w, 4), dtype=np.uint8)
    out[:, :, 3] = 255
    for c in range(nchan):
        out[:, :, c] = sm[c].astype(np.uint8)
    return out


# ---------------------------------------------------------------------------
# Many textures with mip chains: the batched-encode workload.

def _uniform(seed, texture, k):
    """Deterministic uniform double in [0, 1) for (seed, texture, k)."""
    x = _splitmix(np.array([seed], np.uint64))[0] ^ _splitmix(np.array([(texture << 8) + k + 0x77], np.uint64))[0]
    v = _splitmix(np.array([x], np.uint64))[0]
    return float(v >> np.uint64(11)) / float(1 << 53)


def batch_sizes(n=300, seed=0x5EED0003, lo=16, hi=4096, ragged=0.30):
    """n (w, h) pairs, each side log-uniform in [lo, hi]; a `ragged` share of the
    textures has a side that is not a multiple of 4, the rest have both sides
    rounded down to multiples of 4."""
    out = []
    span = np.log2(hi / lo)
    for t in range(n):
        w = min(hi, max(lo, int(lo * 2.0 ** (span * _uniform(seed, t, 0)))))
        h = min(hi, max(lo, int(lo * 2.0 ** (span * _uniform(seed, t, 1)))))
        if _uniform(seed, t, 2) < ragged:
            if w % 4 == 0 and h % 4 == 0:
                d = 1 + int(_uniform(seed, t, 3) * 3)
                w = w + d if w + d <= hi else w - d
        else:
            w, h = w - w % 4, h - h % 4
        out.append((w, h))
    return out


def alpha_channel(w, h, seed, texture=0, binary=0.15):
    """(h, w) uint8: uniform noise smoothed by a 3x3 box ((sum + 4) // 9, edge
    clamp); a `binary` share of the 4x4 blocks instead holds texels of 0 or 255."""
    words = _block_seeds(seed, texture, w * h)
    vals = (words & np.uint64(0xFF)).astype(np.int32).reshape(h, w)
    pad = np.pad(vals, 1, mode='edge')
    acc = np.zeros_like(vals)
    for dy in range(3):
        for dx in range(3):
            acc += pad[dy:dy + h, dx:dx + w]
    sm = (acc + 4) // 9
    bw, bh = (w + 3) // 4, (h + 3) // 4
    bseeds = _block_seeds(seed ^ 0xA1FA, texture, bw * bh)
    on = ((bseeds % np.uint64(100)) < np.uint64(int(round(binary * 100)))).reshape(bh, bw)
    on = np.repeat(np.repeat(on, 4, 0), 4, 1)[:h, :w]
    bit = ((words >> np.uint64(8)) & np.uint64(1)).astype(np.int32).reshape(h, w) * 255
    return np.where(on, bit, sm).astype(np.uint8)


def mip_sizes(w, h):
    """Mip chain sizes: floor halving down to 1x1 (quicktex/image_utils.py:9-36)."""
    out = [(w, h)]
    while out[-1] != (1, 1):
        pw, ph = out[-1]
        out.append((max(pw // 2, 1), max(ph // 2, 1)))
    return out


def box_halve(img, nw, nh):
    """Next mip: 2x2 box (a + b + c + d + 2) >> 2, edge clamp on odd sides."""
    a = img.astype(np.int32)
    h, w = a.shape[:2]
    ys = np.minimum(np.arange(nh) * 2, h - 1)
    xs = np.minimum(np.arange(nw) * 2, w - 1)
    ys1, xs1 = np.minimum(ys + 1, h - 1), np.minimum(xs + 1, w - 1)
    return ((a[ys][:, xs] + a[ys][:, xs1] + a[ys1][:, xs] + a[ys1][:, xs1] + 2) >> 2).astype(np.uint8)


def pad4(img):
    """Pad to multiples of 4 with edge-clamped texels."""
    h, w = img.shape[:2]
    ph, pw = (-h) % 4, (-w) % 4
    if ph == 0 and pw == 0:
        return img
    return np.pad(img, ((0, ph), (0, pw), (0, 0)), mode='edge')


def batch_textures(n=300, seed=0x5EED0003):
    """The batched-encode workload: n RGBA8 textures of batch_sizes(), RGB of the
    block mixture and alpha_channel(), each with its full box-filtered mip chain.

    Returns a list of (texture, level, w, h, rgba) where w x h is the unpadded mip
    size and rgba the mip padded to multiples of 4 by edge clamp, which is what is
    encoded."""
    sizes = batch_sizes(n, seed)

    def one(t):
        w, h = sizes[t]
        img = mixture_rgba(w, h, seed=seed, texture=t, threads=False)
        img[..., 3] = alpha_channel(w, h, seed ^ 0x5A5A, texture=t)
        out = []
        for lvl, (mw, mh) in enumerate(mip_sizes(w, h)):
            if lvl:
                img = box_halve(img, mw, mh)
            out.append((t, lvl, mw, mh, pad4(img)))
        return out

    res = []
    for chain in _pool().map(one, range(n)):
        res.extend(chain)
    return res
