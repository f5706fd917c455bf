"""GPU parity at the benchmark configurations' full sizes.

* BC1 level 18, ThreeColorBlack, 4096 x 4096 (1 M blocks: the default automatic
  choice takes the queued lane kernel over every CU and parks endpoint searches
  for the drain kernel), each interpolator: every block against the 16-lane
  encoder's bytes, and the first, the last and every 16th block row against the
  compiled reference.
* BC4 (channel 0) and BC5 (channels 0, 1) encode of a 2048 x 2048 texture of
  smoothed noise with 20% per-block constants, and the decode of those blocks:
  every byte against the reference.
* BC3 level 10 of 300 textures of log-uniform sides 16..4096 with their full mip
  chains (2925 jobs) as one batch: every job against its plain launch, and the
  chains of every 10th texture against the reference.
* A serving run with Poisson arrivals of mixed BC1/BC3/BC4/BC5 encode and
  decode requests (sides log-uniform 32..1024, levels 0..18): every result
  against the same request served alone by the plain host entry point.

Textures are the seeded synthetic inputs of quicktex_amd.synth, the same the
benchmark times.
"""
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

from quicktex_amd import synth

pytestmark = pytest.mark.gpu


@pytest.fixture(scope='module')
def reference_or_oracle():
    from tests._cpu_libs import oracle, reference
    return reference() or oracle()


@pytest.fixture(scope='module')
def c2_image():
    return synth.mixture_rgba(4096, 4096, seed=0x5EED0002, near_black=0.10)


def _first_bad(a, b):
    bad = np.nonzero((a != b).reshape(a.shape[0], -1).any(axis=1))[0]
    return len(bad), bad[:4].tolist()


@pytest.mark.parametrize('interp', [0, 1, 2, 3])
def test_c2_bc1_level18_4096(qtx, reference_or_oracle, c2_image, interp):
    W = H = 4096
    p = qtx.params(18, 0x13, interp)
    print('C2 interpolator %d' % interp, flush=True)
    auto = qtx.bc1_encode(c2_image, p).reshape(H // 4, W // 4 * 8)
    qtx.set_bc1_encoder(2)
    try:
        group = qtx.bc1_encode(c2_image, p).reshape(H // 4, W // 4 * 8)
    finally:
        qtx.set_bc1_encoder(0)
    n, first = _first_bad(auto, group)
    assert n == 0, '%d block rows differ between the queued lane kernel and the 16-lane encoder, first %s' % (n, first)
    rows = sorted(set(range(0, H // 4, 16)) | {H // 4 - 1})
    strip = np.concatenate([c2_image[4 * r:4 * r + 4] for r in rows], axis=0)
    want = reference_or_oracle.bc1_encode(strip, p).reshape(len(rows), W // 4 * 8)
    n, first = _first_bad(auto[rows], want)
    assert n == 0, '%d sampled block rows differ from the reference, first (sample index) %s' % (n, first)


def test_c1_bc4_bc5_2048(qtx, reference_or_oracle):
    W = H = 2048
    img = synth.smooth_channels(W, H, seed=0x5EED0001, nchan=2)
    ref = reference_or_oracle
    b4 = qtx.bc4_encode(img, 0)
    assert (b4 == ref.bc4_encode(img, 0)).all(), 'BC4 encode'
    b5 = qtx.bc5_encode(img, 0, 1)
    assert (b5 == ref.bc5_encode(img, 0, 1)).all(), 'BC5 encode'
    assert (qtx.bc4_decode(b4, W, H, 0) == ref.bc4_decode(b4, W, H, 0)).all(), 'BC4 decode'
    assert (qtx.bc5_decode(b5, W, H, 0, 1) == ref.bc5_decode(b5, W, H, 0, 1)).all(), 'BC5 decode'
    # the round trip keeps every solid block exactly (single-value BC4 blocks)
    dec = qtx.bc5_decode(b5, W, H, 0, 1)
    blocks = img.reshape(H // 4, 4, W // 4, 4, 4)[..., :2]
    solid = (blocks == blocks[:, :1, :, :1]).all(axis=(1, 3)).all(axis=-1)
    back = dec.reshape(H // 4, 4, W // 4, 4, 4)[..., :2]
    assert solid.mean() > 0.15
    assert (back[solid.nonzero()[0], :, solid.nonzero()[1]] == blocks[solid.nonzero()[0], :, solid.nonzero()[1]]).all()


def test_c3_batch_300_textures(qtx, reference_or_oracle):
    from quicktex_amd import hip
    from quicktex_amd.capi import FMT_BC3, OP_ENCODE, Batch, job
    tex = synth.batch_textures(300)
    p = qtx.params(10, 4, 0)
    in_off, out_off, ti, to = [], [], 0, 0
    for (_, _, _, _, img) in tex:
        in_off.append(ti)
        out_off.append(to)
        ti += (img.nbytes + 255) & ~255
        to += ((img.shape[0] // 4) * (img.shape[1] // 4) * 16 + 255) & ~255
    host = np.zeros(ti, np.uint8)
    for (_, _, _, _, img), o in zip(tex, in_off):
        host[o:o + img.nbytes] = img.reshape(-1)
    d_in = hip.Buffer.from_array(host)
    d_plain, d_batch = hip.Buffer(to), hip.Buffer(to)
    d_plain.zero(0x11)
    d_batch.zero(0x22)
    s = hip.Stream()
    dims = [(img.shape[1], img.shape[0]) for (_, _, _, _, img) in tex]
    print('C3: %d jobs' % len(tex), flush=True)
    for (w, h), io, oo in zip(dims, in_off, out_off):
        qtx.bc3_encode_device(p, d_in.ptr + io, w, h, d_plain.ptr + oo, s.handle)
    b = Batch(qtx, [job(OP_ENCODE, FMT_BC3, w, h, d_in.ptr + io, d_batch.ptr + oo, params=p)
                    for (w, h), io, oo in zip(dims, in_off, out_off)])
    b.launch(s.handle)
    s.synchronize()
    plain, batch = d_plain.download(), d_batch.download()
    b.close()
    s.destroy()
    for k, ((w, h), oo) in enumerate(zip(dims, out_off)):
        n = (w // 4) * (h // 4) * 16
        assert (plain[oo:oo + n] == batch[oo:oo + n]).all(), 'job %d (%dx%d) differs from its plain launch' % (k, w, h)
    sample = [k for k, (t, _, _, _, _) in enumerate(tex) if t % 10 == 0]
    with ThreadPoolExecutor(8) as ex:
        want = list(ex.map(lambda k: reference_or_oracle.bc3_encode(tex[k][4], p), sample))
    for k, wnt in zip(sample, want):
        oo = out_off[k]
        assert (batch[oo:oo + wnt.size] == wnt).all(), 'job %d (texture %d level %d) differs from the reference' % (
            k, tex[k][0], tex[k][1])


def test_c4_poisson_serving(qtx):
    from quicktex_amd import hip
    from quicktex_amd.capi import FMT_BC1, FMT_BC3, FMT_BC4, FMT_BC5, OP_DECODE, OP_ENCODE, Server, job
    rng = np.random.default_rng(0x5EED0C4)
    S = 1024
    mix = synth.mixture_rgba(S, S, seed=0x5EED0000)
    mix_a = mix.copy()
    mix_a[..., 3] = synth.alpha_channel(S, S, 0x5EED0005)
    smooth = synth.smooth_channels(S, S, seed=0x5EED0001, nchan=2)
    enc5 = {FMT_BC1: qtx.bc1_encode(mix, qtx.params(5)), FMT_BC3: qtx.bc3_encode(mix_a, qtx.params(5)),
            FMT_BC4: qtx.bc4_encode(smooth, 0), FMT_BC5: qtx.bc5_encode(smooth, 0, 1)}
    reqs = []
    for i in range(240):
        fmt = int(rng.choice([FMT_BC1, FMT_BC3, FMT_BC4, FMT_BC5], p=[0.4, 0.3, 0.15, 0.15]))
        op = OP_ENCODE if rng.random() < 0.5 else OP_DECODE
        w, h = int(32 * 32.0 ** rng.random()), int(32 * 32.0 ** rng.random())
        level = int(rng.integers(0, 19))
        bpb = 8 if fmt in (FMT_BC1, FMT_BC4) else 16
        nb = ((w + 3) // 4) * ((h + 3) // 4)
        if op == OP_ENCODE:
            src = (mix if fmt == FMT_BC1 else mix_a if fmt == FMT_BC3 else smooth)[:h, :w].copy()
            out_n = nb * bpb
        else:
            src = enc5[fmt][:nb * bpb].copy()
            out_n = w * h * 4
        reqs.append(dict(fmt=fmt, op=op, w=w, h=h, p=qtx.params(level, 4, 0), src=src, out_n=out_n))

    def alone(r):
        w, h, s = r['w'], r['h'], r['src']
        if r['op'] == OP_ENCODE:
            f = {FMT_BC1: lambda: qtx.bc1_encode(s, r['p']), FMT_BC3: lambda: qtx.bc3_encode(s, r['p']),
                 FMT_BC4: lambda: qtx.bc4_encode(s, 0), FMT_BC5: lambda: qtx.bc5_encode(s, 0, 1)}[r['fmt']]
        else:
            f = {FMT_BC1: lambda: qtx.bc1_decode(s, w, h, 0, False), FMT_BC3: lambda: qtx.bc3_decode(s, w, h, 0),
                 FMT_BC4: lambda: qtx.bc4_decode(s, w, h, 0), FMT_BC5: lambda: qtx.bc5_decode(s, w, h, 0, 1)}[r['fmt']]
        return np.asarray(f()).reshape(-1)

    want = [alone(r) for r in reqs]
    # half the destinations pinned (written in place), half pageable (staged)
    pinned = [hip.Pinned(r['out_n']) if i % 2 else None for i, r in enumerate(reqs)]
    outs = [pm.array if pm is not None else np.zeros(r['out_n'], np.uint8) for pm, r in zip(pinned, reqs)]
    gaps = rng.exponential(1.0 / 4000.0, len(reqs))
    arrivals = np.cumsum(gaps) - gaps[0]
    srv = Server(qtx, device_bytes=96 << 20, staging_bytes=32 << 20, slots=4, max_batch_jobs=64)
    try:
        base = time.perf_counter()
        tickets = []
        for i, r in enumerate(reqs):
            while time.perf_counter() - base < arrivals[i]:
                pass
            j = job(r['op'], r['fmt'], r['w'], r['h'], 0, 0, params=r['p'], chan0=0, chan1=1, interpolator=0)
            tickets.append(srv.submit(j, np.ascontiguousarray(r['src']).reshape(-1), outs[i]))
        res = [srv.wait(t, 120000) for t in tickets]
        st = srv.stats()
    finally:
        srv.stop()
    for i, (r, x) in enumerate(zip(res, outs)):
        assert r.status == 0, (i, r.message)
        assert (x == want[i]).all(), 'request %d (fmt %d op %d %dx%d) differs from the same request alone' % (
            i, reqs[i]['fmt'], reqs[i]['op'], reqs[i]['w'], reqs[i]['h'])
    assert st['completed'] == len(reqs) and st['batches'] < len(reqs)
    lat = sorted((r.done_s - r.submit_s) * 1000.0 for r in res)
    print('C4: %d requests, %d batches, p50 %.2f ms, p99 %.2f ms' % (
        len(reqs), st['batches'], lat[len(lat) // 2], lat[int(len(lat) * 0.99)]), flush=True)
