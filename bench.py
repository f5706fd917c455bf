#!/usr/bin/env python3
"""Benchmark of the S3TC block-encode hot path on one MI355X.

Workload (the metric's configuration): BC1 level-18 encode of a 4096x4096 RGBA8
texture -- block mixture of 22.5% solid, 22.5% two-colour, 45% noisy gradient
and 10% near-black blocks (quicktex_amd/synth.py) -- in ThreeColorBlack mode, run
once per interpolator (Ideal, IdealRound, Nvidia, AMD).  One step = the four
encodes, inputs already resident in HBM.

Prints one JSON line.  value = Mpixel/s over the timed steps.  roofline: the BC1
encode kernel against the VALU peak (256 CU x 4 SIMD x 16 lanes x 2.4 GHz =
39.3 T lane-ops/s; the encoder is integer/FP32 ALU work, no HBM or MFMA bound),
achieved = algorithmic ops (oracle accounting on a sample, scaled to the launch)
over the kernel's mean duration from HIP events.  cpu_baseline: the reference's
own C++ (oracle/_ref, OpenMP over block rows like the reference) timed on this
host on the first --cpu-rows rows of the same texture; the GPU output of those
rows is checked bit-exact against it.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = 'BC1 L18 encode Mpixel/s bit-exact, batched/served vs plain; p50/p99 latency'
VALU_PEAK_TOPS = 256 * 4 * 16 * 2.4e9 / 1e12  # non-packed 32-bit lane-ops per second, T
W = H = 4096
INTERPS = ('Ideal', 'IdealRound', 'Nvidia', 'AMD')


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_cores():
    n = os.environ.get('OMP_NUM_THREADS')
    return int(n) if n and n.isdigit() else len(os.sched_getaffinity(0))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=5)
    ap.add_argument('--warmup', type=int, default=2)
    ap.add_argument('--cpu-rows', type=int, default=512, help='texture rows timed on the CPU (multiple of 4)')
    ap.add_argument('--no-cpu', action='store_true')
    ap.add_argument('--copy-probe', type=int, default=0, metavar='MIB',
                    help='only run 3 device-to-device copies of MIB MiB (calibrates the HBM byte counters) and exit')
    a = ap.parse_args()
    if a.gpus != 1:
        print('bench.py measures one GPU: the codec has no multi-GPU path (textures shard with no '
              'communication); run with --gpus 1', file=sys.stderr)
        sys.exit(2)

    import numpy as np

    from quicktex_amd import hip, synth
    from quicktex_amd.capi import COLOR_THREE_BLACK, Qtx

    q = Qtx()
    q.ctx  # device set-up through the library's HIP runtime
    if a.copy_probe:
        n = a.copy_probe << 20
        src = hip.Buffer(n)
        src.zero(7)
        dst = hip.Buffer(n)
        for _ in range(3):
            dst.copy_from(src)
        print(json.dumps({'copy_probe_bytes': n, 'copies': 3}), flush=True)
        return

    log('generating %dx%d texture' % (W, H))
    img = synth.mixture_rgba(W, H, seed=0x5EED0002, near_black=0.10)
    d_in = hip.Buffer.from_array(img)
    nb = (W // 4) * (H // 4)
    d_out = hip.Buffer(4 * nb * 8)
    params = [q.params(18, COLOR_THREE_BLACK, i) for i in range(4)]
    stream = hip.Stream()
    sh = stream.handle
    in_ptr, out_ptr = d_in.ptr, d_out.ptr

    def launch(i):
        q.bc1_encode_device(params[i], in_ptr, W, H, out_ptr + i * nb * 8, sh)

    log('warmup %d steps' % a.warmup)
    for _ in range(a.warmup):
        for i in range(4):
            launch(i)
    stream.synchronize()

    # per-launch kernel durations, HIP events on the launch stream
    ev = [[hip.Event() for _ in range(2)] for _ in range(4 * a.steps)]
    t0 = time.perf_counter()
    for s in range(a.steps):
        for i in range(4):
            e = ev[s * 4 + i]
            e[0].record(stream)
            launch(i)
            e[1].record(stream)
    stream.synchronize()
    t1 = time.perf_counter()
    dur = np.array([e[0].elapsed_ms(e[1]) for e in ev]).reshape(a.steps, 4)  # ms
    step_ms = (t1 - t0) * 1000.0 / a.steps
    px_step = 4 * W * H
    value = px_step / (step_ms / 1000.0) / 1e6
    per_interp = {INTERPS[i]: round(W * H / (dur[:, i].mean() / 1000.0) / 1e6, 1) for i in range(4)}
    log('step %.2f ms, %.1f Mpx/s' % (step_ms, value))
    gpu_out = d_out.download().reshape(4, nb, 8)

    # algorithmic ops: oracle accounting on a 16-row sample, scaled per launch
    from tests._cpu_libs import op_count
    sample = np.ascontiguousarray(img[:16])
    ops_block = []
    for i in range(4):
        counts, weights = op_count(lambda o: o.bc1_encode(sample, params[i]))
        ops_block.append(sum(c * w for c, w in zip(counts, weights)) / ((W // 4) * 4))
    ops_launch = [ob * nb for ob in ops_block]
    achieved = [ops_launch[i] / (dur[:, i].mean() / 1000.0) / 1e12 for i in range(4)]
    achieved_mean = float(np.mean(achieved))
    roofline = {
        'kernel': 'k_bc1_encode (level 18, ThreeColorBlack)',
        'bound': 'valu', 'unit': 'TOP/s',
        'achieved': round(achieved_mean, 3), 'peak': round(VALU_PEAK_TOPS, 2),
        'frac': round(achieved_mean / VALU_PEAK_TOPS, 4),
        'algorithmic_ops_per_block': round(float(np.mean(ops_block)), 1),
        'kernel_ms_mean': round(float(dur.mean()), 3),
        'traffic': None,
    }

    cpu = None
    exact = None
    if not a.no_cpu:
        from tests._cpu_libs import oracle, reference
        ref = reference()
        kind = 'reference' if ref is not None else 'port'
        lib = ref if ref is not None else oracle()
        rows = a.cpu_rows - a.cpu_rows % 4
        part = np.ascontiguousarray(img[:rows])
        nbp = (W // 4) * (rows // 4)
        lib.bc1_encode(np.ascontiguousarray(img[:4]), params[0])  # encoder construction, not timed
        secs = 0.0
        exact = True
        for i in range(4):
            t = time.perf_counter()
            cpu_blocks = lib.bc1_encode(part, params[i])
            secs += time.perf_counter() - t
            exact &= bool((cpu_blocks.reshape(nbp, 8) == gpu_out[i, :nbp]).all())
        cpu_mpx = 4 * W * rows / secs / 1e6
        cpu = {'value': round(cpu_mpx, 3), 'unit': 'Mpixel/s', 'cores': cpu_cores(), 'kind': kind,
               'sample': '%d of %d rows (%d blocks) per interpolator, 4 interpolators, %.1f s' % (
                   rows, H, nbp, secs)}
        log('cpu %s: %.3f Mpx/s, bit-exact on sample: %s' % (kind, cpu_mpx, exact))

    line = {
        'metric': METRIC, 'value': round(value, 1), 'unit': 'Mpixel/s', 'n_gpus': 1,
        'steps': a.steps, 'warmup': a.warmup, 'ms_per_step': round(step_ms, 3),
        'higher_is_better': True, 'scaling': 'weak', 'vs_baseline': None,
        'dtype': 'int32/fp32', 'data': 'synthetic',
        'config': {'workload': 'BC1 level-18 encode, 4096x4096 RGBA8, ThreeColorBlack, 4 interpolators per step',
                   'texture': '%dx%d' % (W, H), 'level': 18, 'color_mode': 'ThreeColorBlack',
                   'interpolators': list(INTERPS)},
        'per_interpolator_mpx_s': per_interp,
        'bit_exact_vs_cpu_sample': exact,
        'roofline': roofline,
        'cpu_baseline': cpu,
    }
    if cpu:
        line['gpu_vs_cpu'] = round(value / cpu['value'], 1)
    print(json.dumps(line), flush=True)
    q.close()


if __name__ == '__main__':
    main()
