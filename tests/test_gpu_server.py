"""GPU tests of the serving form (qtx_server_* in include/qtx.h).

Requests of mixed formats, sizes, levels and directions are submitted from
several threads while earlier ones run; each result must equal the plain host
entry point's output for the same request alone (and the CPU checker's on a
subset).  Admission under a small device budget must keep every request
correct while batching them; validation errors come back from submit.
"""
import threading
import time

import numpy as np
import pytest

from quicktex_amd import hip, synth
from quicktex_amd.capi import FMT_BC1, OP_ENCODE, QtxError, Server, job
from tests._cases import Case

pytestmark = pytest.mark.gpu

MIX_SIZES = [(1, 1), (7, 3), (16, 16), (64, 40), (130, 66), (256, 256), (300, 20), (512, 128), (37, 501), (1024, 64)]


@pytest.fixture(scope='module')
def checker():
    from tests._cpu_libs import oracle, reference
    return reference() or oracle()


def make_cases(qtx, n, seed):
    rng = np.random.default_rng(seed)
    return [Case(i, rng, qtx, size=MIX_SIZES[int(rng.integers(0, len(MIX_SIZES)))]) for i in range(n)]


def submit_all(srv, cases, outs, threads=3):
    tickets = [None] * len(cases)

    def worker(k):
        for i in range(k, len(cases), threads):
            c = cases[i]
            j = c.job(0, 0)
            tickets[i] = srv.submit(j, np.ascontiguousarray(c.src).reshape(-1), outs[i])

    ts = [threading.Thread(target=worker, args=(k,)) for k in range(threads)]
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    return tickets


@pytest.mark.parametrize('encoder', [0, 1])  # auto, lane
def test_server_mix_matches_alone(qtx, checker, oracle, encoder):
    cases = make_cases(qtx, 160, 0x5E12)
    alone = [np.asarray(c.host_plain(qtx)).reshape(-1) for c in cases]
    # half the outputs in pinned memory (written directly), half pageable (staged)
    pinned = [hip.Pinned(c.out_bytes) if i % 2 else None for i, c in enumerate(cases)]
    outs = [p.array if p is not None else np.zeros(c.out_bytes, np.uint8) for p, c in zip(pinned, cases)]
    qtx.set_bc1_encoder(encoder)
    srv = Server(qtx, device_bytes=64 << 20, staging_bytes=32 << 20, slots=3, max_batch_jobs=48)
    qtx.set_bc1_encoder(0)
    try:
        tickets = submit_all(srv, cases, outs)
        results = [srv.wait(t, 60000) for t in tickets]
        st = srv.stats()
    finally:
        srv.stop()
    for i, (c, r, o, a) in enumerate(zip(cases, results, outs, alone)):
        assert r.status == 0, (i, r.message)
        assert r.submit_s <= r.admit_s <= r.launch_s <= r.done_s
        assert (o == a).all(), 'request %d (%s fmt %d %dx%d level %d) differs from the plain path' % (
            i, 'enc' if c.op == OP_ENCODE else 'dec', c.fmt, c.w, c.h, c.level)
    for i in range(0, len(cases), 8):
        c = cases[i]
        want = np.asarray(c.cpu(checker, oracle)).reshape(-1)
        assert (outs[i] == want).all(), 'request %d differs from the CPU checker' % i
    assert st['completed'] == len(cases) and st['failed'] == 0
    assert st['batches'] < len(cases)  # requests were batched
    assert st['device_bytes_peak'] <= 64 << 20


def test_server_admission_under_small_budget(qtx):
    # a 4 MiB budget holds only a few requests at a time: the rest wait for memory
    cases = make_cases(qtx, 60, 0x5E13)
    cases = [c for c in cases if (c.src.nbytes + c.out_bytes) < (3 << 20)]
    alone = [np.asarray(c.host_plain(qtx)).reshape(-1) for c in cases]
    outs = [np.zeros(c.out_bytes, np.uint8) for c in cases]
    srv = Server(qtx, device_bytes=4 << 20, staging_bytes=4 << 20, slots=2, max_batch_jobs=8)
    try:
        tickets = submit_all(srv, cases, outs, threads=2)
        for t in tickets:
            assert srv.wait(t, 60000).status == 0
        st = srv.stats()
    finally:
        srv.stop()
    for i, (o, a) in enumerate(zip(outs, alone)):
        assert (o == a).all(), i
    assert st['device_bytes_peak'] <= 4 << 20 and st['max_batch_jobs'] <= 8


def test_server_errors(qtx):
    srv = Server(qtx, device_bytes=1 << 20, staging_bytes=1 << 20)
    try:
        src = np.zeros(64 * 64 * 4, np.uint8)
        dst = np.zeros(64 * 64, np.uint8)
        bad = job(OP_ENCODE, 2, 64, 64, 0, 0)
        with pytest.raises(QtxError) as e:
            srv.submit(bad, src, dst)
        assert e.value.status == 1 and 'unknown format' in str(e.value)
        small = np.zeros(10, np.uint8)
        with pytest.raises(QtxError) as e:
            srv.submit(job(OP_ENCODE, FMT_BC1, 64, 64, 0, 0, params=qtx.params(5)), src, small)
        assert e.value.status == 1 and 'needed' in str(e.value)
        big = np.zeros(1024 * 1024 * 4, np.uint8)
        with pytest.raises(QtxError) as e:
            srv.submit(job(OP_ENCODE, FMT_BC1, 1024, 1024, 0, 0, params=qtx.params(5)), big,
                       np.zeros(512 * 1024, np.uint8))
        assert e.value.status == 3  # ENOMEM: larger than the whole device budget
        with pytest.raises(QtxError) as e:
            srv.wait(12345, 0)
        assert e.value.status == 1
        # a good request still works after the rejected ones
        t = srv.submit(job(OP_ENCODE, FMT_BC1, 64, 64, 0, 0, params=qtx.params(5)), src, np.zeros(2048, np.uint8))
        assert srv.wait(t, 60000).status == 0
    finally:
        srv.stop()


def test_server_stop_drains(qtx):
    cases = make_cases(qtx, 40, 0x5E14)
    outs = [np.zeros(c.out_bytes, np.uint8) for c in cases]
    alone = [np.asarray(c.host_plain(qtx)).reshape(-1) for c in cases]
    srv = Server(qtx, device_bytes=32 << 20, staging_bytes=32 << 20)
    for c, o in zip(cases, outs):
        srv.submit(c.job(0, 0), np.ascontiguousarray(c.src).reshape(-1), o)
    srv.stop()  # returns once every queued request is done
    for o, a in zip(outs, alone):
        assert (o == a).all()


def _bc1_request(qtx, w, h, level, seed):
    img = synth.mixture_rgba(w, h, seed=seed, near_black=0.1)
    p = qtx.params(level, 0x13, 0)
    return img, p, job(OP_ENCODE, FMT_BC1, w, h, 0, 0, params=p)


def test_server_cancel_waiting(qtx):
    """Requests queued behind a long one are cancelled before admission: they end
    with QTX_ECANCELED and untouched buffers; the rest complete as alone."""
    big, pbig, jbig = _bc1_request(qtx, 4096, 2048, 18, 0xCA0)
    smalls = [_bc1_request(qtx, 64, 48, 5, 0xCA1 + i) for i in range(8)]
    srv = Server(qtx, device_bytes=256 << 20, staging_bytes=64 << 20, slots=1, max_batch_jobs=1)
    try:
        outs = [np.full(jbig.width * jbig.height // 2, 0x5A, np.uint8)] + \
               [np.full(16 * 12 * 8, 0x5A, np.uint8) for _ in smalls]
        tickets = [srv.submit(jbig, big.reshape(-1), outs[0])]
        tickets += [srv.submit(j, img.reshape(-1), o) for (img, _, j), o in zip(smalls, outs[1:])]
        cancelled = set()
        for k in (8, 6, 4):
            try:
                srv.cancel(tickets[k])
                cancelled.add(k)
            except QtxError as e:
                assert e.status == 5, e  # EBUSY: already admitted
        res = [srv.wait(t, 120000) for t in tickets]
        st = srv.stats()
    finally:
        srv.stop()
    assert cancelled, 'no request was still waiting behind a 4096x2048 level-18 encode'
    for k, r in enumerate(res):
        if k in cancelled:
            assert r.status == 7 and b'cancelled' in r.message, (k, r.status)  # QTX_ECANCELED
            assert (outs[k] == 0x5A).all(), 'a cancelled request wrote its output'
        else:
            assert r.status == 0, (k, r.message)
            img, p = (big, pbig) if k == 0 else smalls[k - 1][:2]
            assert (outs[k] == qtx.bc1_encode(img, p)).all(), k
    assert st['cancelled'] == len(cancelled) and st['completed'] == len(tickets) - len(cancelled)
    with pytest.raises(QtxError):
        srv2 = Server(qtx, device_bytes=1 << 20, staging_bytes=1 << 20)
        try:
            srv2.cancel(999)  # unknown ticket: EINVAL
        finally:
            srv2.stop()


def test_server_batch_fault_is_contained(qtx, monkeypatch):
    """A batch whose launch fails on the host (forced: the server's fault hook makes
    its third batch fail before any of its work is enqueued) fails its own requests
    only: the requests of the batches in flight beside it and after it are correct,
    and the server takes new requests afterwards."""
    monkeypatch.setenv('QTX_SERVER_INJECT_FAULT', '3')
    reqs = [_bc1_request(qtx, 256, 128, 13, 0xFA0 + i) for i in range(6)]
    srv = Server(qtx, device_bytes=16 << 20, staging_bytes=16 << 20, slots=2, max_batch_jobs=1)
    monkeypatch.delenv('QTX_SERVER_INJECT_FAULT')
    try:
        # all queued at once: one request per batch in arrival order (batch k is
        # request k - 1), two batches in flight at a time
        outs = [np.zeros(64 * 32 * 8, np.uint8) for _ in reqs]
        tickets = [srv.submit(j, img.reshape(-1), o) for (img, _, j), o in zip(reqs, outs)]
        results = [srv.wait(t, 60000) for t in tickets]
        img, p, j = _bc1_request(qtx, 96, 64, 10, 0xFA9)
        late = np.zeros(24 * 16 * 8, np.uint8)
        r_late = srv.wait(srv.submit(j, img.reshape(-1), late), 60000)
        st = srv.stats()
    finally:
        srv.stop()
    for k, (r, out, (im, pk, _)) in enumerate(zip(results, outs, reqs)):
        if k == 2:
            assert r.status == 4 and b'batch enqueue failed' in r.message, (r.status, r.message)
            assert not out.any(), 'the failed request wrote its output'
        else:
            assert r.status == 0, (k, r.message)
            assert (out == qtx.bc1_encode(im, pk)).all(), k
    assert r_late.status == 0 and (late == qtx.bc1_encode(img, p)).all()
    assert st['failed'] == 1 and st['completed'] == 6 and st['batches'] == 7


def test_server_small_request_overtakes_long_batch():
    """With consecutive batches on the server's two streams, a small request admitted
    while a long level-18 batch runs finishes before it.  Run in a fresh process
    (tests/_served_overtake.py), as a serving process would hold the device: the
    runtime maps a process's streams onto a few hardware queues (four by default)
    and shares them beyond that, and this process holds streams of other tests."""
    import json
    import os
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, '-m', 'tests._served_overtake'], cwd=root, capture_output=True, text=True,
                       timeout=180)
    print(r.stdout[-2000:], r.stderr[-2000:], flush=True)
    assert r.returncode == 0, r.stderr[-2000:]
    out = json.loads(r.stdout.strip().splitlines()[-1])
    assert out['status'] == [0, 0] and out['batch_jobs'] == [1, 1], out
    assert out['exact'] == [True, True], out
    assert out['small_done_ms'] < out['big_done_ms'], out
