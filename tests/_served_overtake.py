"""A serving scenario run in a process of its own by test_gpu_server.py: a long
level-18 BC1 batch (4096 x 4096, pinned input read in place), then a small
level-18 request submitted once the long batch is launched.  Prints one JSON
line: both requests' status, batch sizes, completion times and whether their
blocks equal the plain host entry point's."""
import json
import time

import numpy as np

from quicktex_amd import hip, synth
from quicktex_amd.capi import FMT_BC1, OP_ENCODE, Qtx, Server, job


def main():
    q = Qtx()
    big = synth.mixture_rgba(4096, 4096, seed=0x0E0, near_black=0.1)
    small = synth.mixture_rgba(64, 64, seed=0x0E1, near_black=0.1)
    p = q.params(18, 0x13, 0)
    pin_big = hip.Pinned(big.nbytes)
    pin_big.array[:] = big.reshape(-1)
    out_big = hip.Pinned(1024 * 1024 * 8)
    pin_small = hip.Pinned(small.nbytes)
    pin_small.array[:] = small.reshape(-1)
    out_small = hip.Pinned(16 * 16 * 8)
    srv = Server(q, device_bytes=64 << 20, staging_bytes=16 << 20, slots=2, max_batch_jobs=4)
    try:
        t_big = srv.submit(job(OP_ENCODE, FMT_BC1, 4096, 4096, 0, 0, params=p), pin_big.array, out_big.array)
        deadline = time.perf_counter() + 10
        while srv.stats()['batches'] < 1 and time.perf_counter() < deadline:
            time.sleep(0.0001)  # the long batch is admitted
        time.sleep(0.0005)      # and launched
        t_small = srv.submit(job(OP_ENCODE, FMT_BC1, 64, 64, 0, 0, params=p), pin_small.array, out_small.array)
        r_small = srv.wait(t_small, 60000)
        r_big = srv.wait(t_big, 60000)
    finally:
        srv.stop()
    exact = [bool(np.array_equal(out_big.array, q.bc1_encode(big, p))),
             bool(np.array_equal(out_small.array, q.bc1_encode(small, p)))]
    print(json.dumps({'status': [r_big.status, r_small.status], 'batch_jobs': [r_big.batch_jobs, r_small.batch_jobs],
                      'big_launch_ms': r_big.launch_s * 1e3, 'big_done_ms': r_big.done_s * 1e3,
                      'small_submit_ms': r_small.submit_s * 1e3, 'small_done_ms': r_small.done_s * 1e3,
                      'exact': exact}), flush=True)
    q.close()


if __name__ == '__main__':
    main()
