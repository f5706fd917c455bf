#!/usr/bin/env python3
"""Benchmark of the S3TC block-encode hot path on one MI355X.

Headline (value): BC1 level-18 encode of a 4096x4096 RGBA8 texture -- block
mixture of 22.5% solid, 22.5% two-colour, 45% noisy gradient and 10% near-black
blocks (quicktex_amd/synth.py) -- in ThreeColorBlack mode, once per
interpolator (Ideal, IdealRound, Nvidia, AMD), one plain launch per texture.
One step = the four encodes, inputs already resident in HBM.  value =
Mpixel/s over the timed steps.

roofline: that kernel against the VALU peak (256 CU x 4 SIMD x 16 lanes x
2.4 GHz = 39.3 T lane-ops/s; the encoder is integer/FP32 ALU work, no HBM or
MFMA bound); achieved = algorithmic ops (oracle accounting on a sample, scaled
to the launch) over the kernel's mean duration from HIP events on its stream.
traffic = HBM bytes per launch from the FETCH_SIZE / WRITE_SIZE counters of the
committed profile, scaled by a plain copy kernel (see --copy-probe).
cpu_baseline: the reference's own C++ (oracle/_ref, OpenMP over block rows like
the reference) timed on this host on the first --cpu-rows rows of the same
texture; the GPU output of those rows is checked bit-exact against it.

batched: the many-texture workload -- 300 RGBA8 textures of log-uniform sides
16..4096 (30% with a side that is not a multiple of 4, padded to multiples of 4
by edge clamp), each with its full box-filtered mip chain; RGB of the block
mixture, alpha smoothed noise with 15% binary blocks -- BC3 level 10 encoded
(a) with one plain launch per mip and (b) as one batch (HIP graph, autotuned);
outputs must be byte-equal, and a sample is checked against the CPU.

Prints one JSON line.
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
# HBM bytes per k_bc1_encode launch on this workload, from the committed counter
# profile (profiles/bc1_l18_encode_4096.md): FETCH_SIZE x 2.000 + WRITE_SIZE x 1.000
# (copy-kernel calibration), mean over the four interpolators.
TRAFFIC_BYTES = (81.1e6 + 81.1e6 + 83.6e6 + 85.9e6 + 21.0e6 + 21.0e6 + 25.8e6 + 27.5e6) / 4


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_cores():
    n = os.environ.get('OMP_NUM_THREADS')
    return int(n) if n and n.isdigit() else len(os.sched_getaffinity(0))


def cpu_lib():
    from tests._cpu_libs import oracle, reference
    ref = reference()
    return (ref, 'reference') if ref is not None else (oracle(), 'port')


# ---------------------------------------------------------------------------
def plain_leg(a, q, hip, np):
    from quicktex_amd import synth
    from quicktex_amd.capi import COLOR_THREE_BLACK

    log('generating %dx%d texture' % (W, H))
    img = synth.mixture_rgba(W, H, seed=0x5EED0002, near_black=0.10)
    d_in = hip.Buffer.from_array(img)
    nb = (W // 4) * (H // 4)
    d_out = hip.Buffer(4 * nb * 8)
    params = [q.params(18, COLOR_THREE_BLACK, i) for i in range(4)]
    stream = hip.Stream()
    sh = stream.handle

    def launch(i):
        q.bc1_encode_device(params[i], d_in.ptr, W, H, d_out.ptr + i * nb * 8, sh)

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
    value = 4 * W * H / (step_ms / 1000.0) / 1e6
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
        'traffic': round(TRAFFIC_BYTES),
        'traffic_unit': 'HBM bytes per launch (counter profile)',
        'algorithmic_bytes': W * H * 4 + nb * 8,
    }

    cpu = None
    exact = None
    if not a.no_cpu:
        lib, kind = cpu_lib()
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
    stream.destroy()
    return value, step_ms, per_interp, roofline, cpu, exact


# ---------------------------------------------------------------------------
def batched_leg(a, q, hip, np):
    from concurrent.futures import ThreadPoolExecutor

    from quicktex_amd import synth
    from quicktex_amd.capi import COLOR_FOUR, FMT_BC3, OP_ENCODE, Batch, job

    log('generating the batch workload')
    t = time.perf_counter()
    tex = synth.batch_textures(a.batch_textures)
    log('  %d jobs in %.1f s' % (len(tex), time.perf_counter() - t))
    npx = sum(w * h for (_, _, w, h, _) in tex)
    params = q.params(10, COLOR_FOUR, 0)
    in_off, out_off = [], []
    ti = to = 0
    for (_, _, _, _, img) in tex:
        ph, pw = img.shape[:2]
        in_off.append(ti)
        out_off.append(to)
        ti += (img.nbytes + 255) & ~255
        to += (((pw // 4) * (ph // 4) * 16) + 255) & ~255
    host_in = np.zeros(ti, np.uint8)
    for (_, _, _, _, img), o in zip(tex, in_off):
        host_in[o:o + img.nbytes] = img.reshape(-1)
    d_in = hip.Buffer.from_array(host_in)
    d_plain = hip.Buffer(to)
    d_batch = hip.Buffer(to)
    d_plain.zero()
    d_batch.zero()
    stream = hip.Stream()
    sh = stream.handle
    dims = [(img.shape[1], img.shape[0]) for (_, _, _, _, img) in tex]

    def plain_pass():
        for (pw, ph), io, oo in zip(dims, in_off, out_off):
            q.bc3_encode_device(params, d_in.ptr + io, pw, ph, d_plain.ptr + oo, sh)

    def timed(fn, steps):
        e0, e1 = hip.Event(), hip.Event()
        stream.synchronize()
        t0 = time.perf_counter()
        e0.record(stream)
        for _ in range(steps):
            fn()
        e1.record(stream)
        stream.synchronize()
        wall = (time.perf_counter() - t0) * 1000.0 / steps
        return wall, e0.elapsed_ms(e1) / steps

    plain_pass()
    plain_wall, plain_gpu = timed(plain_pass, a.batch_steps)
    log('plain: %.2f ms per pass (%d launches), %.1f Mpx/s' % (plain_wall, len(tex), npx / plain_wall / 1e3))

    jobs = [job(OP_ENCODE, FMT_BC3, pw, ph, d_in.ptr + io, d_batch.ptr + oo, params=params)
            for (pw, ph), io, oo in zip(dims, in_off, out_off)]
    t = time.perf_counter()
    b = Batch(q, jobs)
    create_ms = (time.perf_counter() - t) * 1000.0
    t = time.perf_counter()
    tuning, tuned_ms = b.autotune(sh, repeats=3)
    tune_s = time.perf_counter() - t
    b.launch(sh)
    batch_wall, batch_gpu = timed(lambda: b.launch(sh), a.batch_steps)
    log('batched: %.2f ms per launch, %.1f Mpx/s, tuning %s' % (batch_wall, npx / batch_wall / 1e3, tuning))
    plain_bytes = d_plain.download()
    batch_bytes = d_batch.download()
    equal = bool(np.array_equal(plain_bytes, batch_bytes))
    b.close()

    cpu = None
    if not a.no_cpu:
        lib, kind = cpu_lib()
        # every 10th texture's full chain, up to about 6 Mpixel
        sample = []
        spx = 0
        for k, (ti_, lvl, w, h, img) in enumerate(tex):
            if ti_ % 10 == 0 and spx < 6e6:
                sample.append(k)
                spx += w * h
        lib.bc3_encode(np.ascontiguousarray(tex[sample[0]][4][:4, :4]), params)  # construction, not timed
        t = time.perf_counter()
        outs = [lib.bc3_encode(tex[k][4], params) for k in sample]
        serial = time.perf_counter() - t
        exact = True
        for k, o in zip(sample, outs):
            n = o.size
            exact &= bool(np.array_equal(o.reshape(-1), batch_bytes[out_off[k]:out_off[k] + n]))
        threads = cpu_cores()
        with ThreadPoolExecutor(threads) as ex:
            t = time.perf_counter()
            list(ex.map(lambda k: lib.bc3_encode(tex[k][4], params), sample))
            par = time.perf_counter() - t
        cpu = {'kind': kind, 'serial_mpx_s': round(spx / serial / 1e6, 3),
               'all_cores_mpx_s': round(spx / par / 1e6, 3), 'cores': threads,
               'sample': '%d mips of every 10th texture, %.2f Mpixel; the reference encodes BC3 on one thread '
                         '(Encoder.h:46,58); all-cores runs the sample mips concurrently' % (len(sample), spx / 1e6),
               'bit_exact_vs_batched': exact}
        log('cpu %s: serial %.2f Mpx/s, %d threads %.2f Mpx/s, exact %s' % (
            kind, cpu['serial_mpx_s'], threads, cpu['all_cores_mpx_s'], exact))
    stream.destroy()
    return {
        'workload': 'BC3 level 10 (Ideal), %d textures + full mip chains = %d jobs, %.1f Mpixel (unpadded)' % (
            a.batch_textures, len(tex), npx / 1e6),
        'plain_mpx_s': round(npx / plain_wall / 1e3, 1),
        'plain_ms': round(plain_wall, 3), 'plain_gpu_span_ms': round(plain_gpu, 3),
        'batched_mpx_s': round(npx / batch_wall / 1e3, 1),
        'batched_ms': round(batch_wall, 3), 'batched_gpu_span_ms': round(batch_gpu, 3),
        'batched_vs_plain': round(plain_wall / batch_wall, 2),
        'tuning': {'workgroups_per_cu': tuning[0], 'unit_kops': tuning[1], 'autotune_s': round(tune_s, 2),
                   'best_ms': round(tuned_ms, 3)},
        'create_ms': round(create_ms, 2),
        'bit_exact_batched_vs_plain': equal,
        'cpu': cpu,
    }


# ---------------------------------------------------------------------------
def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=5)
    ap.add_argument('--warmup', type=int, default=2)
    ap.add_argument('--cpu-rows', type=int, default=2048, help='texture rows timed on the CPU (multiple of 4)')
    ap.add_argument('--no-cpu', action='store_true')
    ap.add_argument('--legs', default='plain,batched', help='comma list of: plain, batched')
    ap.add_argument('--batch-textures', type=int, default=300)
    ap.add_argument('--batch-steps', type=int, default=5)
    ap.add_argument('--copy-probe', type=int, default=0, metavar='MIB',
                    help='only run 3 device-to-device copies of MIB MiB (calibrates the HBM byte counters) and exit')
    a = ap.parse_args()
    if a.gpus != 1:
        print('bench.py measures one GPU: the codec has no multi-GPU path (textures shard with no '
              'communication); run with --gpus 1', file=sys.stderr)
        sys.exit(2)
    legs = set(a.legs.split(','))

    import numpy as np

    from quicktex_amd import hip
    from quicktex_amd.capi import Qtx

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

    line = {'metric': METRIC, 'value': None, 'unit': 'Mpixel/s', 'n_gpus': 1, 'steps': a.steps,
            'warmup': a.warmup, 'ms_per_step': None, 'higher_is_better': True, 'scaling': 'weak',
            'vs_baseline': None, 'dtype': 'int32/fp32', 'data': 'synthetic',
            'config': {'workload': 'BC1 level-18 encode, 4096x4096 RGBA8, ThreeColorBlack, 4 interpolators per step',
                       'texture': '%dx%d' % (W, H), 'level': 18, 'color_mode': 'ThreeColorBlack',
                       'interpolators': list(INTERPS)}}
    if 'plain' in legs:
        value, step_ms, per_interp, roofline, cpu, exact = plain_leg(a, q, hip, np)
        line.update({'value': round(value, 1), 'ms_per_step': round(step_ms, 3),
                     'per_interpolator_mpx_s': per_interp, 'bit_exact_vs_cpu_sample': exact,
                     'roofline': roofline, 'cpu_baseline': cpu})
        if cpu:
            line['gpu_vs_cpu'] = round(value / cpu['value'], 1)
    if 'batched' in legs:
        line['batched'] = batched_leg(a, q, hip, np)
    print(json.dumps(line), flush=True)
    q.close()


if __name__ == '__main__':
    main()
