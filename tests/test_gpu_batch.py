"""GPU tests of the batched form (qtx_batch_* in include/qtx.h).

Every job of a batch must produce the bytes that the plain entry point gives
for the same texture alone, and those of the CPU checker (the reference's own
C++ where built, else the oracle): mixed formats, encode and decode, every
interpolator, colour modes and levels 0-19, ragged sizes down to 1x1, and a
full mip chain.  The same batch is relaunched under every launch shape and
after autotuning; each run must give identical bytes.
"""
import ctypes

import numpy as np
import pytest

from quicktex_amd import hip, synth
from quicktex_amd.capi import FMT_BC1, FMT_BC3, FMT_BC4, OP_DECODE, OP_ENCODE, Batch, QtxError, job, run_host_batch
from tests._cases import Case, nblocks

pytestmark = pytest.mark.gpu

@pytest.fixture(scope='module')
def checker():
    from tests._cpu_libs import oracle, reference
    return reference() or oracle()


@pytest.fixture(scope='module')
def stream(qtx):
    qtx.ctx  # the library's context sets the device up first
    s = hip.Stream()
    yield s
    s.destroy()


@pytest.fixture(scope='module')
def mixed(qtx, stream, checker, oracle):
    rng = np.random.default_rng(0xBA7C4)
    cases = [Case(i, rng, qtx) for i in range(48)]
    srcs = [hip.Buffer.from_array(np.ascontiguousarray(c.src).reshape(-1)) for c in cases]
    want = [np.asarray(c.cpu(checker, oracle)).reshape(-1) for c in cases]
    return cases, srcs, want


def _outs(cases):
    outs = [hip.Buffer(c.out_bytes) for c in cases]
    for o in outs:
        o.zero()
    return outs


def _outs_n(sizes):
    outs = [hip.Buffer(n) for n in sizes]
    for o in outs:
        o.zero()
    return outs


def _check(cases, outs, want, what):
    for k, (c, o, w) in enumerate(zip(cases, outs, want)):
        got = o.download()
        assert got.size == w.size
        assert (got == w).all(), '%s: job %d (%s fmt %d %dx%d level %d mode %d interp %d) differs' % (
            what, k, 'enc' if c.op == OP_ENCODE else 'dec', c.fmt, c.w, c.h, c.level, c.mode, c.interp)


@pytest.mark.parametrize('encoder', [0, 1, 2])  # auto, lane, group
def test_batch_matches_alone_and_checker(qtx, stream, mixed, encoder):
    cases, srcs, want = mixed
    qtx.set_bc1_encoder(encoder)
    alone = _outs(cases)
    for c, s, o in zip(cases, srcs, alone):
        c.plain(qtx, s.ptr, o.ptr, stream.handle)
    stream.synchronize()
    _check(cases, alone, want, 'plain')
    outs = _outs(cases)
    b = Batch(qtx, [c.job(s.ptr, o.ptr) for c, s, o in zip(cases, srcs, outs)])
    b.launch(stream.handle)
    stream.synchronize()
    _check(cases, outs, want, 'batch')
    b.close()
    qtx.set_bc1_encoder(0)


def test_batch_relaunch_under_every_tuning(qtx, stream, mixed):
    cases, srcs, want = mixed
    outs = _outs(cases)
    b = Batch(qtx, [c.job(s.ptr, o.ptr) for c, s, o in zip(cases, srcs, outs)])
    for wg in (1, 2, 3, 4):
        for kops in (16, 64, 256, 1024, 16384):
            b.tuning = (wg, kops)
            assert b.tuning == (wg, kops)
            for rep in range(2):  # the work queue must be back at zero for the second run
                for o in outs:
                    o.zero()
                b.launch(stream.handle)
                stream.synchronize()
                _check(cases, outs, want, 'tuning (%d, %d) run %d' % (wg, kops, rep))
    b.close()


def test_batch_autotune(qtx, stream, mixed):
    cases, srcs, want = mixed
    outs = _outs(cases)
    b = Batch(qtx, [c.job(s.ptr, o.ptr) for c, s, o in zip(cases, srcs, outs)])
    (wg, kops), ms = b.autotune(stream.handle, repeats=1)
    assert 1 <= wg <= 4 and kops in (64, 256, 1024, 4096) and ms > 0
    assert b.tuning == (wg, kops)
    for o in outs:
        o.zero()
    b.launch(stream.handle)
    stream.synchronize()
    _check(cases, outs, want, 'autotuned')
    b.close()


def test_batch_run_host(qtx, mixed):
    cases, _, want = mixed
    jobs = []
    outs = []
    for c in cases:
        out = np.zeros(c.out_bytes, np.uint8)
        outs.append(out)
        jobs.append((c.job(0, 0), np.ascontiguousarray(c.src).reshape(-1), out))
    run_host_batch(qtx, jobs)
    for k, (o, w) in enumerate(zip(outs, want)):
        assert (o == w).all(), 'job %d' % k


def mip_chain(img):
    """Box-filtered mip chain with floor halving down to 1x1 (image_utils.py:9-36 sizes)."""
    out = [img]
    while out[-1].shape[0] > 1 or out[-1].shape[1] > 1:
        a = out[-1].astype(np.int32)
        h, w = a.shape[:2]
        nh, nw = max(h // 2, 1), max(w // 2, 1)
        ys = np.minimum(np.arange(nh) * 2, h - 1)
        xs = np.minimum(np.arange(nw) * 2, w - 1)
        ys1 = np.minimum(ys + 1, h - 1)
        xs1 = np.minimum(xs + 1, w - 1)
        m = (a[ys][:, xs] + a[ys][:, xs1] + a[ys1][:, xs] + a[ys1][:, xs1] + 2) >> 2
        out.append(m.astype(np.uint8))
    return out


def test_batch_mip_chain_bc3(qtx, stream, checker):
    img = synth.mixture_rgba(300, 202, seed=0x5EED2000)
    img[..., 3] = synth.smooth_channels(300, 202, seed=0x5EED2001, nchan=1)[..., 0]
    chain = mip_chain(img)
    p = qtx.params(10, 4, 0)
    srcs = [hip.Buffer.from_array(m) for m in chain]
    outs = _outs_n([nblocks(m.shape[1], m.shape[0]) * 16 for m in chain])
    jobs = [job(OP_ENCODE, FMT_BC3, m.shape[1], m.shape[0], s.ptr, o.ptr, params=p)
            for m, s, o in zip(chain, srcs, outs)]
    b = Batch(qtx, jobs)
    b.launch(stream.handle)
    stream.synchronize()
    for m, o in zip(chain, outs):
        assert (o.download() == checker.bc3_encode(m, p)).all(), m.shape
    b.close()


@pytest.mark.parametrize('levels', [(10, 12, 13, 16, 18), (0, 3, 5, 8)])
def test_lane_batch_forms(qtx, stream, checker, levels):
    """BC1/BC3 jobs forced onto the batch's lane kernel, at levels with endpoint
    search and cluster fit (10-18) and without (0-8): ragged sizes, both formats,
    every interpolator and colour mode, launched twice."""
    rng = np.random.default_rng(0x1A4E + levels[0])
    qtx.set_bc1_encoder(1)
    try:
        cases = []
        for k in range(24):
            w, h = [(1, 1), (37, 53), (130, 18), (64, 64), (256, 96), (5, 300)][k % 6]
            fmt = FMT_BC3 if k % 3 == 0 else FMT_BC1
            mode = 4 if fmt == FMT_BC3 else (4, 3, 0x13)[k % 3]
            p = qtx.params(levels[k % len(levels)], mode, k % 4)
            img = synth.mixture_rgba(w, h, seed=int(rng.integers(1 << 30)))
            cases.append((fmt, w, h, p, img))
        srcs = [hip.Buffer.from_array(img) for (_, _, _, _, img) in cases]
        outs = _outs_n([nblocks(w, h) * (16 if fmt == FMT_BC3 else 8) for (fmt, w, h, _, _) in cases])
        jobs = [job(OP_ENCODE, fmt, w, h, s.ptr, o.ptr, params=p) for (fmt, w, h, p, _), s, o in zip(cases, srcs, outs)]
        b = Batch(qtx, jobs)
        for rep in range(2):
            for o in outs:
                o.zero()
            b.launch(stream.handle)
            stream.synchronize()
            for k, ((fmt, w, h, p, img), o) in enumerate(zip(cases, outs)):
                want = checker.bc3_encode(img, p) if fmt == FMT_BC3 else checker.bc1_encode(img, p)
                assert (o.download() == np.asarray(want).reshape(-1)).all(), (rep, k, fmt, w, h)
        b.close()
    finally:
        qtx.set_bc1_encoder(0)


def test_batch_errors(qtx, stream):
    d = hip.Buffer(64 * 64 * 4)
    p3 = qtx.params(5, 3, 0)
    good = job(OP_ENCODE, FMT_BC1, 16, 16, d.ptr, d.ptr, params=qtx.params(5))
    bad = [
        (job(OP_ENCODE, 2, 16, 16, d.ptr, d.ptr), 'unknown format'),
        (job(OP_ENCODE, FMT_BC3, 16, 16, d.ptr, d.ptr, params=p3), 'FourColor'),
        (job(OP_ENCODE, FMT_BC1, 0, 16, d.ptr, d.ptr, params=p3), 'width'),
        (job(OP_ENCODE, FMT_BC1, 16, 16, 0, d.ptr, params=p3), 'null'),
        (job(OP_DECODE, FMT_BC4, 16, 16, d.ptr, d.ptr, chan0=4), 'Channel'),
        (job(OP_DECODE, FMT_BC1, 16, 16, d.ptr, d.ptr, interpolator=7), 'interpolator'),
        (job(5, FMT_BC1, 16, 16, d.ptr, d.ptr), 'op'),
    ]
    for bj, msg in bad:
        with pytest.raises(QtxError) as e:
            Batch(qtx, [good, bj])
        assert e.value.status == 1 and 'job 1' in str(e.value) and msg in str(e.value), str(e.value)
    with pytest.raises(QtxError):
        Batch(qtx, [])
    b = Batch(qtx, [good])
    for t in ((0, 256), (5, 256), (4, 8), (4, 1 << 20)):
        with pytest.raises(QtxError):
            b.tuning = t
    b.close()
    assert ctypes.sizeof(type(good)) == 96


def test_batch_run_host_staged(qtx):
    """qtx_batch_run_host with sources and destinations past the staging threshold:
    large and tiny jobs interleaved, so that parts straddle the 8 MiB chunks both
    ways; every output equals the plain host entry point's."""
    from quicktex_amd import synth
    rng = np.random.default_rng(0x57A6)
    p = qtx.params(6, 4, 2)
    specs = [(OP_ENCODE, FMT_BC1, 1500, 1100), (OP_ENCODE, FMT_BC3, 17, 9), (OP_DECODE, FMT_BC4, 2048, 1100),
             (OP_ENCODE, FMT_BC4, 2100, 1300), (OP_DECODE, FMT_BC1, 5, 3), (OP_ENCODE, FMT_BC3, 1800, 1500),
             (OP_DECODE, FMT_BC3, 1200, 1024)]
    jobs, want = [], []
    for k, (op, fmt, w, h) in enumerate(specs):
        nb = ((w + 3) // 4) * ((h + 3) // 4)
        bpb = 16 if fmt == FMT_BC3 else 8
        if op == OP_ENCODE:
            src = synth.mixture_rgba(w, h, seed=0x57A0 + k, near_black=0.1)
            out = np.zeros(nb * bpb, np.uint8)
            ref = {FMT_BC1: lambda: qtx.bc1_encode(src, p), FMT_BC3: lambda: qtx.bc3_encode(src, p),
                   FMT_BC4: lambda: qtx.bc4_encode(src, 0)}[fmt]()
            j = job(op, fmt, w, h, 0, 0, params=p, chan0=0)
        else:
            src = rng.integers(0, 256, nb * bpb, dtype=np.uint8)
            out = np.zeros(w * h * 4, np.uint8)
            ref = {FMT_BC1: lambda: qtx.bc1_decode(src, w, h, 2, True), FMT_BC3: lambda: qtx.bc3_decode(src, w, h, 2),
                   FMT_BC4: lambda: qtx.bc4_decode(src, w, h, 0)}[fmt]()
            j = job(op, fmt, w, h, 0, 0, chan0=0, write_alpha=True, interpolator=2)
        jobs.append((j, np.ascontiguousarray(src).reshape(-1), out))
        want.append(np.asarray(ref).reshape(-1))
    for _ in range(2):  # twice: the second run reuses the context's buffers
        for _, _, o in jobs:
            o[:] = 0  # as the plain calls' outputs start (a BC4 decode writes its channel only)
        run_host_batch(qtx, jobs)
        for k, ((_, _, o), w) in enumerate(zip(jobs, want)):
            assert (o == w).all(), 'job %d %s differs' % (k, specs[k])
