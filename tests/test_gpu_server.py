"""GPU tests of the serving form (qtx_server_* in include/qtx.h).

Requests of mixed formats, sizes, levels and directions are submitted from
several threads while earlier ones run; each result must equal the plain host
entry point's output for the same request alone (and the CPU checker's on a
subset).  Admission under a small device budget must keep every request
correct while batching them; validation errors come back from submit.
"""
import threading

import numpy as np
import pytest

from quicktex_amd import hip
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
