#!/usr/bin/env python3
"""Benchmark of the S3TC block-encode hot path on one MI355X.

Headline (value): BC1 level-18 encode of a 4096x4096 RGBA8 texture -- block
mixture of 22.5% solid, 22.5% two-colour, 45% noisy gradient and 10% near-black
blocks (quicktex_amd/synth.py) -- in ThreeColorBlack mode, once per
interpolator (Ideal, IdealRound, Nvidia, AMD), one plain encode per texture
(the queued lane kernel, then the drain kernel for the endpoint searches it
parked).  One step = the four encodes, inputs already resident in HBM.  value =
Mpixel/s over the timed steps, after --warmup steps that bring the shader clock
up from idle.  sustained: --sustain-steps more steps back to back with the shader
clock sampled; the last half of those steps gives the sustained rate.

roofline: the encode against the VALU peak (256 CU x 4 SIMD x 16 lanes x
2.4 GHz = 39.3 T lane-ops/s; the encoder is integer/FP32 ALU work, no HBM or
MFMA bound); achieved = algorithmic ops (oracle accounting on a sample, scaled
to the texture) over the encode's mean duration (both kernels) from HIP events
on its stream.  traffic = HBM bytes per encode from the FETCH_SIZE / WRITE_SIZE
counters of the committed profile, scaled by a plain copy kernel (see
--copy-probe).
cpu_baseline: the reference's own C++ (oracle/_ref, OpenMP over block rows like
the reference) timed on this host on the first --cpu-rows rows of the same
texture; the GPU output of those rows is checked bit-exact against it.

batched: the many-texture workload -- 300 RGBA8 textures of log-uniform sides
16..4096 (30% with a side that is not a multiple of 4, padded to multiples of 4
by edge clamp), each with its full box-filtered mip chain; RGB of the block
mixture, alpha smoothed noise with 15% binary blocks -- BC3 level 10 encoded
(a) with one plain launch per mip and (b) as one batch (HIP graph, autotuned);
outputs must be byte-equal, and a sample is checked against the CPU.

served: 2000 requests -- BC1/BC3/BC4/BC5 at 40/30/15/15%, encode:decode 50:50,
sides log-uniform 32..4096, levels 0..18; encode sources of the block mixture
(BC3 with noise alpha) and smoothed channels (BC4/BC5), decode sources their
level-5 encodes -- from pinned host buffers: (a) each request alone through the
plain host entry point, one after another; (b) all submitted at once to the
serving loop; (c) submitted at Poisson arrivals of a fixed rate, for p50/p99
latency.  Outputs of a sample of requests must be byte-equal across (a)-(c).

dds: quicktex.dds.encode's path -- a 4096x4096 RGBA8 image and its full mip
chain, every level resampled from the base with Pillow's BILINEAR filter, BC3
level 10: the resample kernels alone (HBM GB/s), resample + batched encode
device-resident, and host to host; the reference's flow (Pillow resize + its
BC3 encoder per level) timed on a 1024x1024 crop and checked bit-exact.

sweep: every BC1 level (FourColor, Ideal) on the 4096x4096 mixture, one plain
launch each, with its fraction of the VALU roofline, and the HBM-bound kernels
(BC4/BC5 encode, the four decoders) at 4096x4096 with their fraction of HBM.

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
# HBM bytes per BC1 encode on this workload (k_bc1_lane_queue + k_bc1_search_drain),
# from the committed counter profile (profiles/bc1_l18_encode_4096.md): FETCH_SIZE x 2 +
# WRITE_SIZE x 1 (copy-kernel calibration), mean over the four interpolators.
TRAFFIC_BYTES = (223.15e6 + 218.22e6 + 256.40e6 + 224.12e6) / 4


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_cores():
    n = os.environ.get('OMP_NUM_THREADS')
    return int(n) if n and n.isdigit() else len(os.sched_getaffinity(0))


def cpu_model():
    try:
        for line in open('/proc/cpuinfo'):
            if line.startswith('model name'):
                return line.split(':', 1)[1].strip()
    except OSError:
        pass
    return None


class ClockSampler:
    """Samples the GPU's shader clock (amdgpu sysfs: the hwmon freq1_input in Hz, else
    the starred level of pp_dpm_sclk) every `period` seconds on a thread while running."""

    def __init__(self, period=0.01):
        import glob
        import threading
        self.period, self.samples, self._stop = period, [], threading.Event()
        self.paths = sorted(glob.glob('/sys/class/drm/card*/device/hwmon/hwmon*/freq1_input'))
        self.dpm = sorted(glob.glob('/sys/class/drm/card*/device/pp_dpm_sclk'))
        self._t = threading.Thread(target=self._run, daemon=True)

    def read(self):
        for p in self.paths:
            try:
                v = int(open(p).read().strip())
                if v > 0:
                    return v / 1e6
            except (OSError, ValueError):
                pass
        for p in self.dpm:
            try:
                for line in open(p):
                    if line.strip().endswith('*'):
                        return float(line.split(':')[1].strip().rstrip('*').strip().lower().rstrip('mhz'))
            except (OSError, ValueError, IndexError):
                pass
        return None

    def _run(self):
        while not self._stop.is_set():
            v = self.read()
            if v is not None:
                self.samples.append((time.perf_counter(), v))
            self._stop.wait(self.period)

    def __enter__(self):
        self._t.start()
        return self

    def __exit__(self, *exc):
        self._stop.set()
        self._t.join()

    def summary(self, np):
        if not self.samples:
            return None
        v = np.array([x for _, x in self.samples])
        half = v[len(v) // 2:]
        return {'source': 'amdgpu sysfs', 'samples': int(v.size), 'mean_mhz': round(float(v.mean()), 1),
                'first_mhz': round(float(v[0]), 1), 'last_half_mean_mhz': round(float(half.mean()), 1),
                'min_mhz': round(float(v.min()), 1), 'max_mhz': round(float(v.max()), 1)}


def op_count_per_block(images, params, encode='bc1_encode'):
    """Algorithmic ops per block (oracle accounting, single-threaded) over sample images."""
    import numpy as np

    from tests._cpu_libs import op_count
    nblocks = sum(((im.shape[1] + 3) // 4) * ((im.shape[0] + 3) // 4) for im in images)

    def run(o):
        for im in images:
            getattr(o, encode)(np.ascontiguousarray(im), params)
    counts, weights = op_count(run)
    return sum(c * w for c, w in zip(counts, weights)) / nblocks


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

    # the drop-in path first, while the process holds no stream but the library's:
    # host RawTexture bytes (pageable) in, host blocks out, per interpolator through
    # qtx_bc1_encode (staged uploads in row bands, encodes, downloads)
    for i in range(4):  # warm-up: staging, clocks, each interpolator's first launch (code loading)
        host_out = q.bc1_encode(img, params[i])
    reps = 2
    t0 = time.perf_counter()
    for _ in range(reps):
        for i in range(4):
            host_out = q.bc1_encode(img, params[i])
    h2h_s = (time.perf_counter() - t0) / (4 * reps)
    log('host to host: %.2f ms per encode, %.1f Mpx/s' % (h2h_s * 1e3, W * H / h2h_s / 1e6))

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
    log('mean encode per step (ms): ' + ' '.join('%.3f' % m for m in dur.mean(axis=1)))
    gpu_out = d_out.download().reshape(4, nb, 8)

    # sustained load: the same steps back to back for a while longer, the shader
    # clock sampled meanwhile
    sustained = None
    if a.sustain_steps > 0:
        sev = [hip.Event() for _ in range(a.sustain_steps + 1)]
        with ClockSampler() as clocks:
            sev[0].record(stream)
            for s in range(a.sustain_steps):
                for i in range(4):
                    launch(i)
                sev[s + 1].record(stream)
            stream.synchronize()
        per_step = np.array([sev[s].elapsed_ms(sev[s + 1]) for s in range(a.sustain_steps)])
        tail = per_step[a.sustain_steps // 2:]
        sustained = {'steps': a.sustain_steps, 'seconds': round(float(per_step.sum()) / 1000.0, 3),
                     'first_step_ms': round(float(per_step[0]), 3),
                     'last_half_ms_per_step': round(float(tail.mean()), 3),
                     'last_half_mpx_s': round(4 * W * H / (float(tail.mean()) / 1000.0) / 1e6, 1),
                     'gpu_clock': clocks.summary(np)}
        log('sustained: %d steps in %.2f s, last half %.3f ms per step (%.1f Mpx/s)' % (
            a.sustain_steps, sustained['seconds'], sustained['last_half_ms_per_step'], sustained['last_half_mpx_s']))

    host_exact = bool(np.array_equal(host_out.reshape(nb, 8), gpu_out[3]))
    host_to_host = {'mpx_s': round(W * H / h2h_s / 1e6, 1), 'ms_per_encode': round(h2h_s * 1e3, 3),
                    'input': 'pageable host memory (numpy), as RawTexture holds it',
                    'process': 'measured before the bench makes its own stream, as in a process of the '
                               'Python drop-in (more streams than the runtime\'s hardware queues make the '
                               'copy stream share one with the encodes)',
                    'equals_device_resident': host_exact}

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
        'kernel': 'k_bc1_lane_queue + k_bc1_search_drain (BC1 level 18, ThreeColorBlack, one block per lane, '
                  'queued dealing; endpoint searches past round 34 parked and finished by the drain)',
        'bound': 'valu', 'unit': 'TOP/s',
        'achieved': round(achieved_mean, 3), 'peak': round(VALU_PEAK_TOPS, 2),
        'frac': round(achieved_mean / VALU_PEAK_TOPS, 4),
        'algorithmic_ops_per_block': round(float(np.mean(ops_block)), 1),
        'kernel_ms_mean': round(float(dur.mean()), 3),
        'traffic': round(TRAFFIC_BYTES),
        'traffic_unit': 'HBM bytes per encode, both kernels (counter profile)',
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
               'cpu_model': cpu_model(),
               'sample': '%d of %d rows (%d blocks) per interpolator, 4 interpolators, %.1f s' % (
                   rows, H, nbp, secs)}
        log('cpu %s: %.3f Mpx/s, bit-exact on sample: %s' % (kind, cpu_mpx, exact))
    stream.destroy()
    return value, step_ms, per_interp, roofline, cpu, exact, sustained, host_to_host


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
    # roofline of the batch: algorithmic ops per block (oracle accounting on the chains
    # of every 25th texture) x the batch's blocks over the batch's device time
    nblk = sum((pw // 4) * (ph // 4) for pw, ph in dims)
    ops_sample = [img for (ti_, _, _, _, img) in tex if ti_ % 25 == 0]
    ops_block = op_count_per_block(ops_sample, params, 'bc3_encode')
    batch_tops = ops_block * nblk / (batch_gpu / 1e3) / 1e12
    roofline = {'kernel': 'k_lane_batch (BC3 level 10: one block per lane, 256 blocks per workgroup) with '
                          'the batch\'s other group kernels, one graph replay',
                'bound': 'valu', 'unit': 'TOP/s', 'achieved': round(batch_tops, 3), 'peak': round(VALU_PEAK_TOPS, 2),
                'frac': round(batch_tops / VALU_PEAK_TOPS, 4), 'algorithmic_ops_per_block': round(ops_block, 1),
                'blocks': nblk, 'kernel_ms_mean': round(batch_gpu, 3),
                'ops_sample': 'full chains of every 25th texture (%d blocks)' % sum(
                    (im.shape[1] // 4) * (im.shape[0] // 4) for im in ops_sample)}
    log('batched roofline: %.0f ops/block, %.2f TOP/s, %.1f%% of VALU' % (ops_block, batch_tops,
                                                                        100 * batch_tops / VALU_PEAK_TOPS))

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
        cpu = {'kind': kind, 'cpu_model': cpu_model(), 'serial_mpx_s': round(spx / serial / 1e6, 3),
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
        'roofline': roofline,
        'cpu': cpu,
    }


# ---------------------------------------------------------------------------
def served_leg(a, q, hip, np):
    import ctypes

    from quicktex_amd import synth
    from quicktex_amd.capi import (COLOR_FOUR, FMT_BC1, FMT_BC3, FMT_BC4, FMT_BC5, OP_DECODE, OP_ENCODE, Server,
                                   job)

    N = a.serve_requests
    rng = np.random.default_rng(0x5EED0004)
    S = 4096
    log('serving workload: %d requests' % N)
    # request sources live in pinned host memory, as a serving front end's receive buffers would
    mix = synth.mixture_rgba(S, S, seed=0x5EED0000)
    mix_a = mix.copy()
    mix_a[..., 3] = synth.alpha_channel(S, S, 0x5EED0005)
    smooth = synth.smooth_channels(S, S, seed=0x5EED0001, nchan=2)
    masters = {}
    for name, img in (('bc1', mix), ('bc3', mix_a), ('bc45', smooth)):
        pm = hip.Pinned(img.nbytes)
        pm.array[:] = img.reshape(-1)
        masters[name] = pm
    # decode sources: encodes of the same content (level 5)
    d_img = hip.Buffer(S * S * 4)
    blocks = {}
    p5 = q.params(5, COLOR_FOUR, 0)
    nbm = (S // 4) * (S // 4)
    for fmt, src, bpb in ((FMT_BC1, 'bc1', 8), (FMT_BC3, 'bc3', 16), (FMT_BC4, 'bc45', 8), (FMT_BC5, 'bc45', 16)):
        d_img.upload(masters[src].array)
        d_blk = hip.Buffer(nbm * bpb)
        if fmt == FMT_BC1:
            q.bc1_encode_device(p5, d_img.ptr, S, S, d_blk.ptr, 0)
        elif fmt == FMT_BC3:
            q.bc3_encode_device(p5, d_img.ptr, S, S, d_blk.ptr, 0)
        elif fmt == FMT_BC4:
            q.bc4_encode_device(0, d_img.ptr, S, S, d_blk.ptr, 0)
        else:
            q.bc5_encode_device(0, 1, d_img.ptr, S, S, d_blk.ptr, 0)
        hip.synchronize()
        pb = hip.Pinned(nbm * bpb)
        pb.array[:] = d_blk.download()
        blocks[fmt] = pb
    d_img.free()

    reqs = []
    out_total = 0
    for i in range(N):
        fmt = int(rng.choice([FMT_BC1, FMT_BC3, FMT_BC4, FMT_BC5], p=[0.4, 0.3, 0.15, 0.15]))
        op = OP_ENCODE if rng.random() < 0.5 else OP_DECODE
        w = int(32 * 128.0 ** rng.random())
        h = int(32 * 128.0 ** rng.random())
        level = int(rng.integers(0, 19))
        bpb = 8 if fmt in (FMT_BC1, FMT_BC4) else 16
        nb = ((w + 3) // 4) * ((h + 3) // 4)
        if op == OP_ENCODE:
            src = masters['bc1' if fmt == FMT_BC1 else 'bc3' if fmt == FMT_BC3 else 'bc45'].array[:w * h * 4]
            out_n = nb * bpb
        else:
            src = blocks[fmt].array[:nb * bpb]
            out_n = w * h * 4
        reqs.append(dict(fmt=fmt, op=op, w=w, h=h, level=level, src=src, out_off=out_total, out_n=out_n,
                         params=q.params(level, COLOR_FOUR, 0)))
        out_total += (out_n + 255) & ~255
    outbuf = hip.Pinned(out_total)
    npx = sum(r['w'] * r['h'] for r in reqs)
    for r in reqs:
        r['dst'] = outbuf.array[r['out_off']:r['out_off'] + r['out_n']]
    log('  %.1f Mpixel in %d requests, %.2f GB of outputs' % (npx / 1e6, N, out_total / 1e9))

    L = q.lib
    u8p = ctypes.POINTER(ctypes.c_uint8)

    def plain(r):
        s, d = r['src'].ctypes.data_as(u8p), r['dst'].ctypes.data_as(u8p)
        w, h, n = r['w'], r['h'], r['out_n']
        if r['op'] == OP_ENCODE:
            if r['fmt'] == FMT_BC1:
                st = L.qtx_bc1_encode(q.ctx, ctypes.byref(r['params']), s, w, h, d, n)
            elif r['fmt'] == FMT_BC3:
                st = L.qtx_bc3_encode(q.ctx, ctypes.byref(r['params']), s, w, h, d, n)
            elif r['fmt'] == FMT_BC4:
                st = L.qtx_bc4_encode(q.ctx, 0, s, w, h, d, n)
            else:
                st = L.qtx_bc5_encode(q.ctx, 0, 1, s, w, h, d, n)
        else:
            if r['fmt'] == FMT_BC1:
                st = L.qtx_bc1_decode(q.ctx, 0, 0, s, w, h, d, n)
            elif r['fmt'] == FMT_BC3:
                st = L.qtx_bc3_decode(q.ctx, 0, s, w, h, d, n)
            elif r['fmt'] == FMT_BC4:
                st = L.qtx_bc4_decode(q.ctx, 0, s, w, h, d, n)
            else:
                st = L.qtx_bc5_decode(q.ctx, 0, 1, s, w, h, d, n)
        q.check(st)

    # plain: each request alone through the host entry point, one after another
    for r in reqs[:8]:
        plain(r)  # warm-up
    service = np.zeros(N)
    t0 = time.perf_counter()
    for i, r in enumerate(reqs):
        t = time.perf_counter()
        plain(r)
        service[i] = time.perf_counter() - t
    plain_s = time.perf_counter() - t0
    sample = list(range(0, N, 20))
    plain_out = {i: reqs[i]['dst'].copy() for i in sample}
    log('plain: %.3f s, %.1f Mpx/s' % (plain_s, npx / plain_s / 1e6))

    def job_of(r):
        return job(r['op'], r['fmt'], r['w'], r['h'], 0, 0, params=r['params'], chan0=0, chan1=1,
                   interpolator=0)

    def serve(arrivals, pageable=None):
        """pageable: (sources, destinations) in pageable memory instead of the pinned ones."""
        outbuf.array[:] = 0
        srcs = [r['src'] for r in reqs] if pageable is None else pageable[0]
        dsts = [r['dst'] for r in reqs] if pageable is None else pageable[1]
        srv = Server(q, device_bytes=a.serve_device_mib << 20, staging_bytes=(64 if pageable is None else 512) << 20,
                     slots=3, max_batch_jobs=1024)
        try:
            jobs = [job_of(r) for r in reqs]
            base = time.perf_counter()
            tickets = []
            for i, r in enumerate(reqs):
                if arrivals is not None:
                    while time.perf_counter() - base < arrivals[i]:
                        pass
                tickets.append(srv.submit(jobs[i], srcs[i], dsts[i]))
            res = [srv.wait(t, 600000) for t in tickets]
            stats = srv.stats()
        finally:
            srv.stop()
        bad = [i for i, x in enumerate(res) if x.status != 0]
        assert not bad, 'served requests failed: %s' % bad[:5]
        sub = np.array([x.submit_s for x in res])
        done = np.array([x.done_s for x in res])
        span = done.max() - sub.min()
        lat = (done - sub) * 1000.0
        exact = all(np.array_equal(plain_out[i], dsts[i]) for i in sample)
        return span, lat, stats, exact, np.array([x.batch_jobs for x in res])

    sat_span, sat_lat, sat_stats, sat_exact, sat_bj = serve(None)
    sat_rate = N / sat_span
    log('served (all queued at once): %.3f s, %.1f Mpx/s, %d batches' % (sat_span, npx / sat_span / 1e6,
                                                                        sat_stats['batches']))
    # the same requests from and to pageable memory: inputs are copied into the
    # device arena, so admission by device footprint takes part
    pg_master = {k: np.array(v.array) for k, v in masters.items()}
    pg_blocks = {k: np.array(v.array) for k, v in blocks.items()}
    pg_src = []
    for r in reqs:
        if r['op'] == OP_ENCODE:
            m = pg_master['bc1' if r['fmt'] == FMT_BC1 else 'bc3' if r['fmt'] == FMT_BC3 else 'bc45']
            pg_src.append(m[:r['w'] * r['h'] * 4])
        else:
            pg_src.append(pg_blocks[r['fmt']][:r['src'].size])
    pg_dst = [np.zeros(r['out_n'], np.uint8) for r in reqs]
    pg_span, pg_lat, pg_stats, pg_exact, pg_bj = serve(None, (pg_src, pg_dst))
    log('served, pageable buffers: %.3f s, %.1f Mpx/s, device memory peak %.1f MiB' % (
        pg_span, npx / pg_span / 1e6, pg_stats['device_bytes_peak'] / 2**20))
    del pg_master, pg_blocks, pg_src, pg_dst
    lam = a.serve_rate if a.serve_rate > 0 else 0.6 * sat_rate
    gaps = rng.exponential(1.0 / lam, N)
    arrivals = np.cumsum(gaps) - gaps[0]
    po_span, po_lat, po_stats, po_exact, po_bj = serve(arrivals)
    log('served (Poisson %.0f req/s): %.3f s, p50 %.2f ms, p99 %.2f ms' % (
        lam, po_span, np.percentile(po_lat, 50), np.percentile(po_lat, 99)))
    # the plain path under the same arrivals: one request at a time in arrival order
    fin = 0.0
    plain_lat = np.zeros(N)
    for i in range(N):
        fin = max(fin, arrivals[i]) + service[i]
        plain_lat[i] = (fin - arrivals[i]) * 1000.0
    for pm in list(masters.values()) + list(blocks.values()) + [outbuf]:
        pm.free()
    pct = lambda x, p: round(float(np.percentile(x, p)), 3)  # noqa: E731
    # PCIe roofline of the saturated run: every request's input read and output
    # written across the link (kernels read pinned inputs and write outputs in host
    # memory), against PCIe 5.0 x16 (64 GB/s each way)
    host_bytes = sum(r['src'].nbytes + r['out_n'] for r in reqs)
    pcie_gbs = host_bytes / sat_span / 1e9
    roofline = {'bound': 'pcie', 'unit': 'GB/s', 'achieved': round(pcie_gbs, 1), 'peak': 128.0,
                'frac': round(pcie_gbs / 128.0, 4), 'traffic': host_bytes,
                'traffic_unit': 'host bytes read + written over the whole run',
                'note': 'peak: PCIe 5.0 x16, 64 GB/s in each direction'}
    return {
        'workload': ('%d requests: BC1/BC3/BC4/BC5 40/30/15/15%%, encode:decode 50:50, sides log-uniform 32..4096, '
                     'levels 0..18 (FourColor, Ideal), %.1f Mpixel; host buffers pinned' % (N, npx / 1e6)),
        'plain_mpx_s': round(npx / plain_s / 1e6, 1),
        'served_mpx_s': round(npx / sat_span / 1e6, 1),
        'served_vs_plain': round(plain_s / sat_span, 2),
        'served_batches': sat_stats['batches'], 'served_mean_batch': round(float(sat_bj.mean()), 1),
        'poisson': {'rate_req_s': round(lam, 1), 'requests': N,
                    'throughput_mpx_s': round(npx / po_span / 1e6, 1),
                    'p50_ms': pct(po_lat, 50), 'p99_ms': pct(po_lat, 99), 'mean_ms': round(float(po_lat.mean()), 3),
                    'batches': po_stats['batches'], 'mean_batch': round(float(po_bj.mean()), 1),
                    'plain_same_arrivals_p50_ms': pct(plain_lat, 50),
                    'plain_same_arrivals_p99_ms': pct(plain_lat, 99)},
        'pageable': {'served_mpx_s': round(npx / pg_span / 1e6, 1), 'span_s': round(pg_span, 3),
                     'batches': pg_stats['batches'], 'p50_ms': pct(pg_lat, 50), 'p99_ms': pct(pg_lat, 99),
                     'device_bytes_peak': int(pg_stats['device_bytes_peak']),
                     'note': 'the same requests all queued at once, sources and destinations in pageable memory: '
                             'inputs are copied into the device arena (admission by device footprint), outputs '
                             'staged through pinned memory'},
        'device_bytes_peak': int(max(sat_stats['device_bytes_peak'], po_stats['device_bytes_peak'],
                                     pg_stats['device_bytes_peak'])),
        'device_budget_bytes': a.serve_device_mib << 20,
        'roofline': roofline,
        'bit_exact_served_vs_plain_sample': bool(sat_exact and po_exact and pg_exact),
    }


# ---------------------------------------------------------------------------
def _events_ms(hip, stream, fn, reps):
    """Mean device time of fn() over reps runs on stream (HIP events), after one untimed run."""
    fn()
    e0, e1 = hip.Event(), hip.Event()
    e0.record(stream)
    for _ in range(reps):
        fn()
    e1.record(stream)
    stream.synchronize()
    return e0.elapsed_ms(e1) / reps


def _median_s(fn, runs=5, warm=1):
    for _ in range(warm):  # warm-up (encoder construction and tables are not timed)
        fn()
    t = []
    for _ in range(runs):
        t0 = time.perf_counter()
        fn()
        t.append(time.perf_counter() - t0)
    return float(sorted(t)[len(t) // 2])


def c0_leg(a, q, hip, np):
    """The reference's CPU-runnable case: BC1 level 5, Ideal, FourColor, one 256x256
    texture of the block mixture (alpha 255)."""
    from quicktex_amd import synth
    S = 256
    img = synth.mixture_rgba(S, S, seed=0x5EED0000)
    p = q.params(5, 4, 0)
    nb = (S // 4) * (S // 4)
    d_in, d_out = hip.Buffer.from_array(img), hip.Buffer(nb * 8)
    stream = hip.Stream()
    ms = _events_ms(hip, stream, lambda: q.bc1_encode_device(p, d_in.ptr, S, S, d_out.ptr, stream.handle), 200)
    gpu = d_out.download()
    # a call takes tens of microseconds: warm the clocks with a few hundred first
    host_s = _median_s(lambda: q.bc1_encode(img, p), 200, warm=200)
    ops_block = op_count_per_block([img], p)
    tops = ops_block * nb / (ms / 1e3) / 1e12
    out = {'workload': 'BC1 level 5 (FourColor, Ideal), 256x256 RGBA8 block mixture, one plain launch',
           'device_mpx_s': round(S * S / ms / 1e3, 1), 'kernel_ms': round(ms, 4),
           'host_to_host_mpx_s': round(S * S / host_s / 1e6, 1), 'host_to_host_ms': round(host_s * 1e3, 3),
           'roofline': {'kernel': 'k_bc1_encode (16 lanes per block; 4096 blocks are below the lane '
                                  'kernels\' crossover)', 'bound': 'valu', 'unit': 'TOP/s',
                        'achieved': round(tops, 4), 'peak': round(VALU_PEAK_TOPS, 2),
                        'frac': round(tops / VALU_PEAK_TOPS, 5), 'algorithmic_ops_per_block': round(ops_block, 1),
                        'note': 'a launch of 4096 blocks fills under 2% of the GPU\'s lanes; its time is one '
                                'block\'s latency'}}
    if not a.no_cpu:
        lib, kind = cpu_lib()
        secs = _median_s(lambda: lib.bc1_encode(img, p), 5)
        exact = bool(np.array_equal(lib.bc1_encode(img, p), gpu))
        out['cpu_baseline'] = {'value': round(S * S / secs / 1e6, 3), 'unit': 'Mpixel/s', 'cores': cpu_cores(),
                               'kind': kind, 'cpu_model': cpu_model(),
                               'sample': 'the whole texture, median of 5 encodes (OpenMP over block rows, as '
                                         'the reference)', 'bit_exact_vs_gpu': exact}
        log('c0: GPU %.4f ms (%.0f Mpx/s), host to host %.3f ms, CPU %.2f Mpx/s, exact %s' % (
            ms, S * S / ms / 1e3, host_s * 1e3, S * S / secs / 1e6, exact))
    stream.destroy()
    return out


# The last-level cache (256 MB of Infinity Cache) holds a texture of tens of MiB
# from one launch to the next, so launches repeated over one buffer read part of
# it from the cache, not from HBM (measured: BC4 encode of 4096x4096 0.0115 ms on
# one buffer, 0.0132 ms rotating over five).  The HBM-bound legs rotate over
# copies whose inputs and outputs together exceed the cache.
ROTATE_BYTES = 768 << 20


def _rotation(per_copy_bytes):
    """Copies to rotate over so that they hold at least ROTATE_BYTES."""
    return max(2, -(-ROTATE_BYTES // per_copy_bytes))


def _cycler(fns):
    """One callable that runs fns[0], fns[1], ... in turn."""
    it = [0]

    def step():
        fns[it[0] % len(fns)]()
        it[0] += 1
    return step


def c1_leg(a, q, hip, np):
    """BC4 (channel 0) and BC5 (channels 0, 1) encode and decode of a 2048x2048
    texture of smoothed noise with 20% per-block constants; each timed launch
    reads a different copy of its input, so that the inputs come from HBM."""
    from quicktex_amd import synth
    S = 2048
    img = synth.smooth_channels(S, S, seed=0x5EED0001, nchan=2)
    nb = (S // 4) * (S // 4)
    K = _rotation(S * S * 4 * 2 + nb * 16)
    d_img = [hip.Buffer.from_array(img) for _ in range(K)]
    d_blk = [hip.Buffer(nb * 16) for _ in range(K)]
    d_out = [hip.Buffer(S * S * 4) for _ in range(K)]
    stream = hip.Stream()
    sh = stream.handle

    def per_copy(call):
        return _cycler([lambda k=k: call(k) for k in range(K)])
    enc4 = lambda k: q.bc4_encode_device(0, d_img[k].ptr, S, S, d_blk[k].ptr, sh)  # noqa: E731
    enc5 = lambda k: q.bc5_encode_device(0, 1, d_img[k].ptr, S, S, d_blk[k].ptr, sh)  # noqa: E731
    cases = [('bc4_encode', 4.5, enc4, None),
             ('bc4_decode', 4.5, lambda k: q.bc4_decode_device(0, d_blk[k].ptr, S, S, d_out[k].ptr, sh), enc4),
             ('bc5_encode', 5.0, enc5, None),
             ('bc5_decode', 5.0, lambda k: q.bc5_decode_device(0, 1, d_blk[k].ptr, S, S, d_out[k].ptr, sh), enc5)]
    out = {'workload': '2048x2048, smoothed noise + 20%% per-block constants; BC4 channel 0, BC5 channels 0, 1',
           'rotation': '%d copies of input and output (%d MiB), one per timed launch in turn' % (
               K, K * (S * S * 8 + nb * 16) >> 20)}
    lib, kind = cpu_lib() if not a.no_cpu else (None, None)
    for name, bpp, call, producer in cases:
        if producer is not None:  # each copy's blocks from its own encode
            for k in range(K):
                producer(k)
        ms = _events_ms(hip, stream, per_copy(call), 4 * K)
        gbs = S * S * bpp / (ms / 1e3) / 1e9
        r = {'kernel_ms': round(ms, 4), 'device_mpx_s': round(S * S / ms / 1e3, 1),
             'roofline': {'bound': 'hbm', 'unit': 'GB/s', 'achieved': round(gbs, 1), 'peak': 8000.0,
                          'frac': round(gbs / 8000.0, 4), 'algorithmic_bytes': int(S * S * bpp)}}
        if lib is not None:
            if name == 'bc4_encode':
                cfn = lambda: lib.bc4_encode(img, 0)  # noqa: E731
            elif name == 'bc5_encode':
                cfn = lambda: lib.bc5_encode(img, 0, 1)  # noqa: E731
            elif name == 'bc4_decode':
                blocks0 = d_blk[0].download(nb * 8)
                cfn = lambda: lib.bc4_decode(blocks0, S, S, 0)  # noqa: E731
            else:
                blocks0 = d_blk[0].download(nb * 16)
                cfn = lambda: lib.bc5_decode(blocks0, S, S, 0, 1)  # noqa: E731
            call(0)
            stream.synchronize()
            got = d_blk[0].download(nb * (8 if '4' in name else 16)) if 'encode' in name else d_out[0].download()
            secs = _median_s(cfn, 5)
            want = cfn()
            r['cpu_baseline'] = {'value': round(S * S / secs / 1e6, 3), 'unit': 'Mpixel/s', 'cores': 1,
                                 'kind': kind, 'cpu_model': cpu_model(),
                                 'sample': 'the whole texture, median of 5 (serial, as the reference runs it)',
                                 'bit_exact_vs_gpu': bool(np.array_equal(np.asarray(want).reshape(-1), got.reshape(-1)))}
        out[name] = r
        log('c1 %s: %.4f ms, %.0f GB/s (%.1f%% of HBM)%s' % (
            name, ms, gbs, gbs / 80.0, '' if lib is None else ', CPU %.1f Mpx/s, exact %s' % (
                r['cpu_baseline']['value'], r['cpu_baseline']['bit_exact_vs_gpu'])))
    stream.destroy()
    return out


def sweep_leg(a, q, hip, np):
    """Plain-kernel throughput per BC1 level (FourColor, Ideal) and of the
    HBM-bound kernels (BC4/BC5 encode, every decoder) on 4096x4096 inputs."""
    from quicktex_amd import synth
    from quicktex_amd.capi import COLOR_FOUR
    from tests._cpu_libs import op_count

    S = a.sweep_size
    img = synth.mixture_rgba(S, S, seed=0x5EED0002, near_black=0.10)
    sm = synth.smooth_channels(S, S, seed=0x5EED0001, nchan=2)
    nb = (S // 4) * (S // 4)
    d_img = hip.Buffer.from_array(img)
    d_sm = hip.Buffer.from_array(sm)
    d_blk = hip.Buffer(nb * 16)
    d_out = hip.Buffer(S * S * 4)
    stream = hip.Stream()
    sh = stream.handle
    sample = np.ascontiguousarray(img[:16])

    def timeit(fn, reps=3):
        fn()
        e0, e1 = hip.Event(), hip.Event()
        e0.record(stream)
        for _ in range(reps):
            fn()
        e1.record(stream)
        stream.synchronize()
        return e0.elapsed_ms(e1) / reps

    levels = []
    for lvl in range(20):
        p = q.params(lvl, COLOR_FOUR, 0)
        ms = timeit(lambda: q.bc1_encode_device(p, d_img.ptr, S, S, d_blk.ptr, sh))
        counts, weights = op_count(lambda o: o.bc1_encode(sample, p))
        ops = sum(c * w for c, w in zip(counts, weights)) / ((S // 4) * 4)
        tops = ops * nb / (ms / 1e3) / 1e12
        levels.append({'level': lvl, 'ms': round(ms, 4), 'mpx_s': round(S * S / ms / 1e3, 1),
                       'ops_per_block': round(ops), 'valu_frac': round(tops / VALU_PEAK_TOPS, 4),
                       'hbm_frac': round((S * S * 4.5) / (ms / 1e3) / 8e12, 4)})
        log('level %2d: %.3f ms, %.0f Mpx/s, %.0f ops/block, VALU %.1f%%' % (
            lvl, ms, S * S / ms / 1e3, ops, 100 * tops / VALU_PEAK_TOPS))
    hbm = {}
    # each timed launch reads a different copy of its input (and writes a different
    # output), so that the inputs come from HBM and not from the last-level cache
    K = _rotation(S * S * 4 * 2 + nb * 16)
    d_sms = [d_sm] + [hip.Buffer.from_array(sm) for _ in range(K - 1)]
    d_blks = [d_blk] + [hip.Buffer(nb * 16) for _ in range(K - 1)]
    d_outs = [d_out] + [hip.Buffer(S * S * 4) for _ in range(K - 1)]
    enc4 = lambda k: q.bc4_encode_device(0, d_sms[k].ptr, S, S, d_blks[k].ptr, sh)  # noqa: E731
    enc5 = lambda k: q.bc5_encode_device(0, 1, d_sms[k].ptr, S, S, d_blks[k].ptr, sh)  # noqa: E731
    bc1_blocks = lambda k: q.bc1_encode_device(q.params(5, COLOR_FOUR, 0), d_sms[k].ptr, S, S,  # noqa: E731
                                               d_blks[k].ptr, sh)
    bc3_blocks = lambda k: q.bc3_encode_device(q.params(5, COLOR_FOUR, 0), d_sms[k].ptr, S, S,  # noqa: E731
                                               d_blks[k].ptr, sh)
    cases = [('bc4_encode', 4.5, enc4, None),
             ('bc5_encode', 5.0, enc5, None),
             ('bc1_decode', 4.5, lambda k: q.bc1_decode_device(0, 0, d_blks[k].ptr, S, S, d_outs[k].ptr, sh),
              bc1_blocks),
             ('bc3_decode', 5.0, lambda k: q.bc3_decode_device(0, d_blks[k].ptr, S, S, d_outs[k].ptr, sh),
              bc3_blocks),
             ('bc4_decode', 4.5, lambda k: q.bc4_decode_device(0, d_blks[k].ptr, S, S, d_outs[k].ptr, sh), enc4),
             ('bc5_decode', 5.0, lambda k: q.bc5_decode_device(0, 1, d_blks[k].ptr, S, S, d_outs[k].ptr, sh),
              enc5)]
    for name, bpp, call, producer in cases:
        if producer is not None:  # each copy's blocks from an encode of its own
            for k in range(K):
                producer(k)
        ms = timeit(_cycler([lambda k=k: call(k) for k in range(K)]), 4 * K)
        gbs = S * S * bpp / (ms / 1e3) / 1e9
        hbm[name] = {'ms': round(ms, 4), 'mpx_s': round(S * S / ms / 1e3, 1), 'gb_s': round(gbs, 1),
                     'hbm_frac': round(gbs / 8000.0, 4)}
        log('%s: %.4f ms, %.0f GB/s (%.1f%% of 8 TB/s)' % (name, ms, gbs, gbs / 80.0))
    hbm['rotation'] = '%d copies of input and output (%d MiB), one per timed launch in turn' % (
        K, K * (S * S * 8 + nb * 16) >> 20)
    stream.destroy()
    return {'size': '%dx%d' % (S, S), 'bc1_levels_fourcolor_ideal': levels, 'hbm_bound': hbm}


# ---------------------------------------------------------------------------
def dds_leg(a, q, hip, np):
    """The DDS encoder's path (quicktex.dds.encode): a mip chain resampled from the
    base with Pillow's BILINEAR filter, every level BC3 level-10 encoded."""
    import ctypes

    from quicktex_amd import synth
    from quicktex_amd.capi import COLOR_FOUR, FMT_BC3, OP_ENCODE, Batch, job, mip_sizes, mips_encode_host

    S = a.dds_size
    img = synth.mixture_rgba(S, S, seed=0x5EED0005)
    img[..., 3] = synth.alpha_channel(S, S, seed=0x5EED0006)
    levels = q.lib.qtx_mip_count(S, S, 0)
    sizes = mip_sizes(S, S, levels)
    npx = sum(w * h for w, h in sizes)
    params = q.params(10, COLOR_FOUR, 0)
    stream = hip.Stream()
    sh = stream.handle

    # (1) the resample kernels alone, device-resident
    d_img = hip.Buffer.from_array(img)
    d_levels = [hip.Buffer(w * h * 4) for w, h in sizes[1:]]
    ws_bytes = q.lib.qtx_mips_workspace_bytes(S, S, levels)
    d_ws = hip.Buffer(ws_bytes)
    ptrs = (ctypes.c_void_p * (levels - 1))(*[b.ptr for b in d_levels])

    def mips():
        q.check(q.lib.qtx_mips_device(q.ctx, d_img.ptr, S, S, levels, ptrs, d_ws.ptr, ws_bytes, sh))

    def events(fn, reps):
        fn()
        e0, e1 = hip.Event(), hip.Event()
        e0.record(stream)
        for _ in range(reps):
            fn()
        e1.record(stream)
        stream.synchronize()
        return e0.elapsed_ms(e1) / reps

    mips_ms = events(mips, 10)
    # each level reads the base once and writes itself
    alg_bytes = sum(S * S * 4 + w * h * 4 for w, h in sizes[1:])
    log('dds: mip resample %.3f ms for %d levels, %.0f GB/s algorithmic' % (mips_ms, levels - 1,
                                                                         alg_bytes / mips_ms / 1e6))

    # (2) device-resident chain encode: resample + one batch over the levels
    d_blocks = [hip.Buffer(((w + 3) // 4) * ((h + 3) // 4) * 16) for w, h in sizes]
    srcs = [d_img.ptr] + [b.ptr for b in d_levels]
    b = Batch(q, [job(OP_ENCODE, FMT_BC3, w, h, s, d.ptr, params=params)
                  for (w, h), s, d in zip(sizes, srcs, d_blocks)])

    def chain():
        mips()
        b.launch(sh)

    chain_ms = events(chain, 5)
    log('dds: device chain (resample + batch encode) %.3f ms, %.1f Mpx/s' % (chain_ms, npx / chain_ms / 1e3))

    # (3) host to host through qtx_mips_encode_host, as quicktex_amd.dds.encode calls it
    t = job(OP_ENCODE, FMT_BC3, 0, 0, 0, 0, params=params)
    mips_encode_host(q, t, img, levels)
    reps = 3
    t0 = time.perf_counter()
    for _ in range(reps):
        host = mips_encode_host(q, t, img, levels)
    host_ms = (time.perf_counter() - t0) * 1000.0 / reps
    log('dds: host to host %.2f ms, %.1f Mpx/s' % (host_ms, npx / host_ms / 1e3))
    same = all(np.array_equal(h, np.frombuffer(d.download(), np.uint8)[:h.size]) for h, d in zip(host, d_blocks))
    b.close()

    cpu = None
    if not a.no_cpu:
        # the reference's flow on a crop: Pillow resize of each level from the base,
        # then the reference's BC3 encoder per level (one thread, Encoder.h:46,58)
        from PIL import Image
        lib, kind = cpu_lib()
        C = a.dds_cpu_size
        crop = np.ascontiguousarray(img[:C, :C])
        im = Image.frombytes('RGBA', (C, C), crop.tobytes())
        csizes = mip_sizes(C, C, q.lib.qtx_mip_count(C, C, 0))
        lib.bc3_encode(np.ascontiguousarray(crop[:4, :4]), params)  # construction, not timed
        t0 = time.perf_counter()
        outs = []
        for m, (w, h) in enumerate(csizes):
            if m == 0:
                lvl = crop
            else:
                rgb = im.convert('RGB').resize((w, h), Image.BILINEAR)
                rgb.putalpha(im.getchannel('A').resize((w, h), Image.BILINEAR))
                lvl = np.asarray(rgb)
            outs.append(lib.bc3_encode(lvl, params))
        secs = time.perf_counter() - t0
        cpx = sum(w * h for w, h in csizes)
        gpu = mips_encode_host(q, t, crop, len(csizes))
        exact = all(np.array_equal(g, o.reshape(-1)) for g, o in zip(gpu, outs))
        cpu = {'kind': kind, 'value': round(cpx / secs / 1e6, 3), 'unit': 'Mpixel/s', 'cores': 1,
               'sample': '%dx%d crop of the same image, full chain (%d levels), Pillow resize + BC3 level 10 '
                         'per level' % (C, C, len(csizes)),
               'bit_exact_vs_gpu': exact}
        log('dds cpu %s: %.3f Mpx/s, exact %s' % (kind, cpu['value'], exact))
    stream.destroy()
    return {
        'workload': 'DDS encode: %dx%d RGBA8 + full mip chain (%d levels, Pillow BILINEAR from the base), '
                    'BC3 level 10, %.1f Mpixel' % (S, S, levels, npx / 1e6),
        'mips_ms': round(mips_ms, 4),
        'mips_algorithmic_gb_s': round(alg_bytes / mips_ms / 1e6, 1),
        'mips_hbm_frac': round(alg_bytes / mips_ms / 1e6 / 8000.0, 4),
        'device_chain_ms': round(chain_ms, 3), 'device_chain_mpx_s': round(npx / chain_ms / 1e3, 1),
        'host_to_host_ms': round(host_ms, 2), 'host_to_host_mpx_s': round(npx / host_ms / 1e3, 1),
        'host_equals_device': bool(same),
        'cpu': cpu,
    }


# ---------------------------------------------------------------------------
def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=20)
    ap.add_argument('--sustain-steps', type=int, default=200,
                    help='plain leg: steps run back to back after the timed ones (0: none)')
    ap.add_argument('--warmup', type=int, default=10,
                    help='untimed steps first (the shader clock ramps up from idle over the first ~50 ms)')
    ap.add_argument('--cpu-rows', type=int, default=2048, help='texture rows timed on the CPU (multiple of 4)')
    ap.add_argument('--no-cpu', action='store_true')
    ap.add_argument('--legs', default='plain,c0,c1,batched,served,dds,sweep',
                    help='comma list of: plain, c0, c1, batched, served, dds, sweep')
    ap.add_argument('--serve-requests', type=int, default=2000)
    ap.add_argument('--serve-rate', type=float, default=0.0,
                    help='Poisson arrival rate, requests/s (default: 60%% of the measured saturation rate)')
    ap.add_argument('--serve-device-mib', type=int, default=4096, help='device budget of the server')
    ap.add_argument('--sweep-size', type=int, default=4096, help='texture side of the sweep leg')
    ap.add_argument('--batch-textures', type=int, default=300)
    ap.add_argument('--batch-steps', type=int, default=5)
    ap.add_argument('--dds-size', type=int, default=4096, help='base side of the dds leg')
    ap.add_argument('--dds-cpu-size', type=int, default=1024, help='base side of the dds leg\'s CPU sample')
    ap.add_argument('--copy-probe', type=int, default=0, metavar='MIB',
                    help='only run 3 device-to-device copies of MIB MiB (calibrates the HBM byte counters) and exit')
    ap.add_argument('--level-probe', type=int, default=-1, metavar='LEVEL',
                    help='only run 3 BC1 encodes of the 4096x4096 texture at LEVEL (FourColor, Ideal; for counter runs) and exit')
    ap.add_argument('--host-probe', type=int, default=-1, metavar='LEVEL',
                    help='only run BC1 encodes of the 4096x4096 texture at LEVEL from pageable host memory through '
                         'qtx_bc1_encode (ThreeColorBlack at 18, else FourColor; Ideal), in a process holding no other '
                         'streams, print their median time and exit')
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
    if a.host_probe >= 0:
        from quicktex_amd import synth
        from quicktex_amd.capi import COLOR_FOUR, COLOR_THREE_BLACK
        img = synth.mixture_rgba(W, H, seed=0x5EED0002, near_black=0.10)
        p = q.params(a.host_probe, COLOR_THREE_BLACK if a.host_probe == 18 else COLOR_FOUR, 0)
        q.bc1_encode(img, p)  # warm-up: staging and band resources, clocks
        t = []
        for _ in range(7):
            t0 = time.perf_counter()
            q.bc1_encode(img, p)
            t.append(time.perf_counter() - t0)
        ms = sorted(t)[3] * 1e3
        print(json.dumps({'host_probe': a.host_probe, 'encodes': 7, 'median_ms': round(ms, 3),
                          'mpx_s': round(W * H / ms / 1e3, 1)}), flush=True)
        return
    if a.level_probe >= 0:
        from quicktex_amd import synth
        from quicktex_amd.capi import COLOR_FOUR
        img = synth.mixture_rgba(W, H, seed=0x5EED0002, near_black=0.10)
        d_in = hip.Buffer.from_array(img)
        d_out = hip.Buffer((W // 4) * (H // 4) * 8)
        stream = hip.Stream()
        p = q.params(a.level_probe, COLOR_FOUR, 0)
        for _ in range(3):
            q.bc1_encode_device(p, d_in.ptr, W, H, d_out.ptr, stream.handle)
        stream.synchronize()
        print(json.dumps({'level_probe': a.level_probe, 'encodes': 3}), flush=True)
        return

    line = {'metric': METRIC, 'value': None, 'unit': 'Mpixel/s', 'n_gpus': 1, 'steps': a.steps,
            'warmup': a.warmup, 'ms_per_step': None, 'higher_is_better': True, 'scaling': 'weak',
            'vs_baseline': None, 'dtype': 'int32/fp32', 'data': 'synthetic',
            'config': {'workload': 'BC1 level-18 encode, 4096x4096 RGBA8, ThreeColorBlack, 4 interpolators per step',
                       'texture': '%dx%d' % (W, H), 'level': 18, 'color_mode': 'ThreeColorBlack',
                       'interpolators': list(INTERPS)}}
    if 'plain' in legs:
        value, step_ms, per_interp, roofline, cpu, exact, sustained, h2h = plain_leg(a, q, hip, np)
        line.update({'value': round(value, 1), 'ms_per_step': round(step_ms, 3),
                     'per_interpolator_mpx_s': per_interp, 'bit_exact_vs_cpu_sample': exact,
                     'roofline': roofline, 'cpu_baseline': cpu, 'sustained': sustained, 'host_to_host': h2h})
        if cpu:
            line['gpu_vs_cpu'] = round(value / cpu['value'], 1)
    if 'c0' in legs:
        line['c0'] = c0_leg(a, q, hip, np)
    if 'c1' in legs:
        line['c1'] = c1_leg(a, q, hip, np)
    if 'batched' in legs:
        line['batched'] = batched_leg(a, q, hip, np)
    if 'served' in legs:
        line['served'] = served_leg(a, q, hip, np)
    if 'dds' in legs:
        line['dds'] = dds_leg(a, q, hip, np)
    if 'sweep' in legs:
        line['sweep'] = sweep_leg(a, q, hip, np)
    print(json.dumps(line), flush=True)
    q.close()


if __name__ == '__main__':
    main()
