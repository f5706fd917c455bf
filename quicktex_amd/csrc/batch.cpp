// batch.cpp -- the batched form of the codec (include/qtx.h, qtx_batch_*).
//
// A batch is a fixed list of jobs (one texture each, any size, format and
// encoder settings) launched as one HIP graph:
//   * BC1/BC3 encodes with endpoint search run in one persistent kernel per
//     interpolator present (k_color_batch); its waves take work units -- runs of
//     consecutive blocks of one job -- from a queue in device memory, so textures
//     of any size share the whole GPU and no block waits for a launch of its own;
//   * BC1/BC3 encodes without endpoint search run one block per lane
//     (k_lane_batch), a workgroup per 256 blocks;
//   * BC4/BC5 encodes and every decode run in one kernel with a workgroup per
//     256 blocks (k_light_batch).
// The kernels are independent branches of the graph.  Jobs are ordered by
// estimated cost, largest first, so the queue ends on small units.  The launch
// shape (resident workgroups per CU, work per unit) is a tuning that
// qtx_batch_autotune picks by timing.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <vector>

#include "kernels.h"
#include "qtx_internal.h"
#include "batch_internal.h"
#include "runtime.h"

using namespace qtx;
using namespace qtx::rt;

namespace qtx {
namespace bt {

// Failure of job i, or of a single request when i is kNoIndex: the message is
// prefixed with "job <i>: " for batches.
static qtx_status jfail(size_t i, qtx_status st, const char *fmt, ...) __attribute__((format(printf, 3, 4)));
static qtx_status jfail(size_t i, qtx_status st, const char *fmt, ...) {
    char buf[400];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (i == kNoIndex) return fail(st, "%s", buf);
    return fail(st, "job %zu: %s", i, buf);
}

qtx_status prepare(const qtx_job &j, size_t i, Prepared &p) {
    std::memset(&p, 0, sizeof p);
    qtx_status st = check_dims(j.width, j.height);
    if (st != QTX_OK) return jfail(i, st, "%s", last_error());
    if (!j.src || !j.dst) return jfail(i, QTX_EINVAL, "null buffer");
    BatchJob &b = p.job;
    b.src = static_cast<const uint8_t *>(j.src);
    b.dst = static_cast<uint8_t *>(j.dst);
    b.w = j.width;
    b.h = j.height;
    b.bw = (j.width + 3) / 4;
    b.bh = (j.height + 3) / 4;
    b.nblocks = (uint32_t)b.bw * (uint32_t)b.bh;
    const int bpb = (j.format == QTX_FMT_BC1 || j.format == QTX_FMT_BC4) ? 8 : 16;
    if (j.format != QTX_FMT_BC1 && j.format != QTX_FMT_BC3 && j.format != QTX_FMT_BC4 && j.format != QTX_FMT_BC5)
        return jfail(i, QTX_EINVAL, "unknown format %d", j.format);
    if (j.op == QTX_OP_ENCODE) {
        p.src_bytes = raw_bytes(j.width, j.height);
        p.dst_bytes = block_bytes(j.width, j.height, bpb);
        if (j.format == QTX_FMT_BC1 || j.format == QTX_FMT_BC3) {
            int interp = 0;
            if (to_settings(&j.bc1, b.s, interp) != QTX_OK) return jfail(i, QTX_EINVAL, "%s", last_error());
            if (j.format == QTX_FMT_BC3 && b.s.color_mode != QTX_COLOR_FOUR)
                return jfail(i, QTX_EINVAL, "BC3 colour blocks are always FourColor");
            b.kind = j.format == QTX_FMT_BC1 ? BJ_ENC_BC1 : BJ_ENC_BC3;
            b.interp = interp;
            p.group = kGroupColour + interp;  // plan() moves it to the lane encoder when that is faster
            p.block_ops = qtx::bc1_block_ops(b.s) + (j.format == QTX_FMT_BC3 ? 300.0 : 0.0);
        } else {
            if (j.chan0 < 0 || j.chan0 > 3 || (j.format == QTX_FMT_BC5 && (j.chan1 < 0 || j.chan1 > 3)))
                return jfail(i, QTX_EINVAL, "Channel out of range");
            b.kind = j.format == QTX_FMT_BC4 ? BJ_ENC_BC4 : BJ_ENC_BC5;
            b.p0 = j.chan0;
            b.p1 = j.format == QTX_FMT_BC5 ? j.chan1 : 0;
            p.group = kGroupLight;
            p.block_ops = j.format == QTX_FMT_BC4 ? 300.0 : 600.0;
        }
    } else if (j.op == QTX_OP_DECODE) {
        p.src_bytes = block_bytes(j.width, j.height, bpb);
        p.dst_bytes = raw_bytes(j.width, j.height);
        p.group = kGroupLight;
        p.block_ops = 200.0;
        switch (j.format) {
            case QTX_FMT_BC1:
            case QTX_FMT_BC3:
                if (j.bc1.interpolator < 0 || j.bc1.interpolator > 3)
                    return jfail(i, QTX_EINVAL, "Invalid interpolator type");
                b.kind = j.format == QTX_FMT_BC1 ? BJ_DEC_BC1 : BJ_DEC_BC3;
                b.interp = j.bc1.interpolator;
                b.p0 = j.format == QTX_FMT_BC1 ? (j.write_alpha != 0) : 0;
                break;
            default:
                if (j.chan0 < 0 || j.chan0 > 3 || (j.format == QTX_FMT_BC5 && (j.chan1 < 0 || j.chan1 > 3)))
                    return jfail(i, QTX_EINVAL, "Channel out of range");
                b.kind = j.format == QTX_FMT_BC4 ? BJ_DEC_BC4 : BJ_DEC_BC5;
                b.p0 = j.chan0;
                b.p1 = j.format == QTX_FMT_BC5 ? j.chan1 : 0;
                break;
        }
    } else {
        return jfail(i, QTX_EINVAL, "unknown op %d", j.op);
    }
    return QTX_OK;
}

// Unit layout of a list of prepared jobs: groups (colour by interpolator, then
// light), each in descending total cost, unit sizes for the tuning.
qtx_status plan(const std::vector<Prepared> &jobs, const qtx_batch_tuning &t, int cu_count, int bc1_path, Plan &out) {
    out.host.clear();
    out.groups.clear();
    // BC1/BC3 jobs of an interpolator go to the lane encoder when there is enough
    // such work in the launch (use_lane_batch with the group's total blocks)
    uint64_t colour_blocks[4] = {0, 0, 0, 0};
    for (const Prepared &p : jobs)
        if (p.group >= kGroupColour && p.group < kGroupLane) colour_blocks[p.group - kGroupColour] += p.job.nblocks;
    std::vector<int> group(jobs.size());
    for (size_t i = 0; i < jobs.size(); i++) {
        const int g = jobs[i].group;
        group[i] = g;
        if (g >= kGroupColour && g < kGroupLane && use_lane_batch(jobs[i].job.s, colour_blocks[g - kGroupColour], bc1_path))
            group[i] = kGroupLane + (g - kGroupColour);
    }
    std::vector<int> idx(jobs.size());
    std::iota(idx.begin(), idx.end(), 0);
    auto total = [&](int i) { return jobs[i].block_ops * jobs[i].job.nblocks; };
    std::stable_sort(idx.begin(), idx.end(), [&](int x, int y) { return total(x) > total(y); });
    for (int grp = 0; grp < kGroups; grp++) {
        Plan::Group g;
        g.kind = grp < kGroupLane ? kGroupColour : (grp < kGroupLight ? kGroupLane : kGroupLight);
        g.interp = grp - g.kind;
        g.first = out.host.size();
        uint64_t unit = 0;
        for (int i : idx) {
            if (group[i] != grp) continue;
            BatchJob j = jobs[i].job;
            if (g.kind == kGroupColour) {
                const double rounds = std::round((double)t.unit_kops * 1000.0 / jobs[i].block_ops);
                j.unit_blocks = 4u * (uint32_t)std::min<double>(std::max(rounds, 1.0), (double)kMaxRounds);
            } else {
                j.unit_blocks = g.kind == kGroupLane ? kLaneUnitBlocks : kLightUnitBlocks;
            }
            j.unit_begin = (uint32_t)unit;
            unit += (j.nblocks + j.unit_blocks - 1) / j.unit_blocks;
            if (unit >= (1ull << 31)) return fail(QTX_ERANGE, "batch has more than 2^31 work units");
            out.host.push_back(j);
            g.order.push_back(i);
        }
        if (g.order.empty()) continue;
        g.nunits = (uint32_t)unit;
        if (g.kind == kGroupColour) {
            const long long waves_needed = ((long long)g.nunits + 3) / 4;  // four waves per workgroup
            g.grid = (unsigned)std::max(1LL, std::min<long long>((long long)cu_count * t.workgroups_per_cu,
                                                                 waves_needed));
        } else {
            g.grid = g.nunits;
        }
        out.groups.push_back(std::move(g));
    }
    return QTX_OK;
}

hipError_t enqueue_group(const Plan::Group &g, const BatchJob *d_jobs, uint32_t *d_queue,
                         const DeviceTables *d_tables, hipStream_t s) {
    const BatchJob *dj = d_jobs + g.first;
    if (g.kind == kGroupColour)
        return launch_color_batch(g.interp, dj, (int)g.order.size(), g.nunits, d_queue, g.grid, d_tables, s);
    if (g.kind == kGroupLane) return launch_lane_batch(g.interp, dj, (int)g.order.size(), g.nunits, g.grid, d_tables, s);
    return launch_light_batch(dj, (int)g.order.size(), g.nunits, g.grid, s);
}

qtx_status run_once(qtx_ctx *ctx, const std::vector<qtx_job> &jobs, hipStream_t s,
                    std::vector<BatchJob> &host_table) {
    std::vector<Prepared> prep(jobs.size());
    qtx_status st;
    for (size_t i = 0; i < jobs.size(); i++)
        if ((st = prepare(jobs[i], i, prep[i])) != QTX_OK) return st;
    Plan pl;
    if ((st = plan(prep, qtx_batch_tuning{kMaxColourWgPerCu, 256}, ctx->cu_count, ctx->bc1_path, pl)) != QTX_OK)
        return st;
    constexpr size_t kQueueBytes = 256;  // two words per kernel group
    static_assert(2 * kGroups * sizeof(uint32_t) <= kQueueBytes, "queue words fit");
    const size_t need = kQueueBytes + pl.host.size() * sizeof(BatchJob);
    if (need > ctx->once_cap) {
        // hipFree waits for the device: the last run's kernels are done with the table
        if (ctx->d_once) (void)hipFree(ctx->d_once);
        ctx->d_once = nullptr;
        ctx->once_cap = 0;
        const size_t cap = std::max(need, (size_t)64 << 10);
        QTX_HIP(hipMalloc(&ctx->d_once, cap));
        const hipError_t e = hipMemset(ctx->d_once, 0, kQueueBytes);
        if (e != hipSuccess) {
            (void)hipFree(ctx->d_once);
            ctx->d_once = nullptr;
            return hip_fail(e, "batch queue words");
        }
        ctx->once_cap = cap;
    }
    uint32_t *queues = reinterpret_cast<uint32_t *>(ctx->d_once);
    auto *table = reinterpret_cast<BatchJob *>(ctx->d_once + kQueueBytes);
    // the table is copied in stream order, after the last run's kernels on s, from
    // the caller's vector
    host_table = std::move(pl.host);
    QTX_HIP(hipMemcpyAsync(table, host_table.data(), host_table.size() * sizeof(BatchJob), hipMemcpyHostToDevice, s));
    for (size_t g = 0; g < pl.groups.size(); g++)
        QTX_HIP(enqueue_group(pl.groups[g], table, queues + 2 * g, ctx->d_tables, s));
    return QTX_OK;
}

}  // namespace bt
}  // namespace qtx

using namespace qtx::bt;

struct qtx_batch {
    qtx_ctx *ctx = nullptr;
    std::vector<Prepared> jobs;
    Plan plan;  // colour groups (by interpolator), then the light group
    BatchJob *d_jobs = nullptr;
    uint32_t *d_queues = nullptr;  // two words per group
    qtx_batch_tuning tuning{kMaxColourWgPerCu, 256};
    int bc1_path = QTX_BC1_AUTO;
    hipStream_t cap = nullptr;
    std::vector<hipStream_t> side;
    std::vector<hipEvent_t> joins;
    hipEvent_t fork = nullptr, done = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

namespace {

void drop_graph(qtx_batch *b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
}

// Unit layout for the tuning, upload of the job array, capture of the graph.
qtx_status rebuild(qtx_batch *b) {
    qtx_status st = plan(b->jobs, b->tuning, b->ctx->cu_count, b->bc1_path, b->plan);
    if (st != QTX_OK) return st;
    const std::vector<BatchJob> &host = b->plan.host;
    if (b->done) QTX_HIP(hipEventSynchronize(b->done));
    QTX_HIP(hipMemcpy(b->d_jobs, host.data(), host.size() * sizeof(BatchJob), hipMemcpyHostToDevice));
    drop_graph(b);

    QTX_HIP(hipStreamBeginCapture(b->cap, hipStreamCaptureModeThreadLocal));
    hipError_t e = hipEventRecord(b->fork, b->cap);
    const auto &groups = b->plan.groups;
    for (size_t i = 0; i < groups.size() && e == hipSuccess; i++) {
        hipStream_t s = i == 0 ? b->cap : b->side[i - 1];
        if (i > 0) e = hipStreamWaitEvent(s, b->fork, 0);
        if (e != hipSuccess) break;
        e = enqueue_group(groups[i], b->d_jobs, b->d_queues + 2 * i, b->ctx->d_tables, s);
        if (e == hipSuccess && i > 0) e = hipEventRecord(b->joins[i - 1], s);
    }
    for (size_t i = 1; i < groups.size() && e == hipSuccess; i++) e = hipStreamWaitEvent(b->cap, b->joins[i - 1], 0);
    hipGraph_t graph = nullptr;
    const hipError_t e2 = hipStreamEndCapture(b->cap, &graph);
    if (e != hipSuccess) {
        if (graph) (void)hipGraphDestroy(graph);
        return hip_fail(e, "batch capture");
    }
    if (e2 != hipSuccess) return hip_fail(e2, "hipStreamEndCapture");
    b->graph = graph;
    QTX_HIP(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    return QTX_OK;
}

qtx_status check_tuning(const qtx_batch_tuning *t) {
    if (!t) return fail(QTX_EINVAL, "null tuning");
    if (t->workgroups_per_cu < 1 || t->workgroups_per_cu > kMaxColourWgPerCu)
        return fail(QTX_EINVAL, "workgroups_per_cu must be in 1..%d", kMaxColourWgPerCu);
    if (t->unit_kops < 16 || t->unit_kops > 16384) return fail(QTX_EINVAL, "unit_kops must be in 16..16384");
    return QTX_OK;
}

}  // namespace

extern "C" {

qtx_status qtx_batch_destroy(qtx_batch *b) {
    if (!b) return QTX_OK;
    (void)hipSetDevice(b->ctx->device);
    if (b->done) (void)hipEventSynchronize(b->done);
    drop_graph(b);
    for (hipEvent_t ev : b->joins) (void)hipEventDestroy(ev);
    for (hipStream_t s : b->side) (void)hipStreamDestroy(s);
    if (b->fork) (void)hipEventDestroy(b->fork);
    if (b->done) (void)hipEventDestroy(b->done);
    if (b->cap) (void)hipStreamDestroy(b->cap);
    if (b->d_jobs) (void)hipFree(b->d_jobs);
    if (b->d_queues) (void)hipFree(b->d_queues);
    delete b;
    return QTX_OK;
}

qtx_status qtx_batch_create(qtx_ctx *ctx, const qtx_job *jobs, size_t njobs, qtx_batch **out) {
    if (!out) return fail(QTX_EINVAL, "null output pointer");
    *out = nullptr;
    qtx_status st = check_ctx(ctx);
    if (st != QTX_OK) return st;
    if (!jobs || njobs == 0) return fail(QTX_EINVAL, "empty batch");
    if (njobs > (1u << 24)) return fail(QTX_ERANGE, "batch of %zu jobs exceeds 2^24", njobs);
    auto *b = new qtx_batch;
    b->ctx = ctx;
    b->bc1_path = ctx->bc1_path;
    b->jobs.resize(njobs);
    for (size_t i = 0; i < njobs; i++) {
        if ((st = prepare(jobs[i], i, b->jobs[i])) != QTX_OK) {
            delete b;
            return st;
        }
    }
    if ((st = plan(b->jobs, b->tuning, ctx->cu_count, b->bc1_path, b->plan)) != QTX_OK) {
        delete b;
        return st;
    }
    const size_t ngroups = b->plan.groups.size();
    hipError_t e = hipStreamCreateWithFlags(&b->cap, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&b->fork, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&b->done, hipEventDisableTiming);
    for (size_t i = 1; i < ngroups && e == hipSuccess; i++) {
        hipStream_t s = nullptr;
        hipEvent_t ev = nullptr;
        e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
        if (e == hipSuccess) b->side.push_back(s);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
        if (e == hipSuccess) b->joins.push_back(ev);
    }
    if (e == hipSuccess) e = hipMalloc(&b->d_jobs, njobs * sizeof(BatchJob));
    if (e == hipSuccess) e = hipMalloc(&b->d_queues, 2 * ngroups * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMemset(b->d_queues, 0, 2 * ngroups * sizeof(uint32_t));
    if (e != hipSuccess) {
        qtx_batch_destroy(b);
        return hip_fail(e, "batch set-up");
    }
    if ((st = rebuild(b)) != QTX_OK) {
        qtx_batch_destroy(b);
        return st;
    }
    *out = b;
    return QTX_OK;
}

qtx_status qtx_batch_launch(qtx_batch *b, void *stream) {
    if (!b) return fail(QTX_EINVAL, "null batch");
    QTX_HIP(hipSetDevice(b->ctx->device));
    hipStream_t s = (hipStream_t)stream;
    // the work queues belong to the batch: runs of one batch never overlap
    QTX_HIP(hipStreamWaitEvent(s, b->done, 0));
    QTX_HIP(hipGraphLaunch(b->exec, s));
    QTX_HIP(hipEventRecord(b->done, s));
    return QTX_OK;
}

qtx_status qtx_batch_set_tuning(qtx_batch *b, const qtx_batch_tuning *t) {
    if (!b) return fail(QTX_EINVAL, "null batch");
    qtx_status st = check_tuning(t);
    if (st != QTX_OK) return st;
    QTX_HIP(hipSetDevice(b->ctx->device));
    b->tuning = *t;
    return rebuild(b);
}

qtx_status qtx_batch_get_tuning(const qtx_batch *b, qtx_batch_tuning *t) {
    if (!b || !t) return fail(QTX_EINVAL, "null argument");
    *t = b->tuning;
    return QTX_OK;
}

qtx_status qtx_batch_autotune(qtx_batch *b, void *stream, int32_t repeats, qtx_batch_tuning *best, float *best_ms) {
    if (!b) return fail(QTX_EINVAL, "null batch");
    if (repeats < 1) return fail(QTX_EINVAL, "repeats must be >= 1");
    QTX_HIP(hipSetDevice(b->ctx->device));
    hipStream_t s = (hipStream_t)stream;
    static const int wgs[] = {2, 3, 4};
    static const uint32_t kops[] = {64, 256, 1024, 4096};
    hipEvent_t t0 = nullptr, t1 = nullptr;
    QTX_HIP(hipEventCreate(&t0));
    hipError_t e = hipEventCreate(&t1);
    if (e != hipSuccess) {
        (void)hipEventDestroy(t0);
        return hip_fail(e, "hipEventCreate");
    }
    qtx_batch_tuning chosen = b->tuning;
    float chosen_ms = INFINITY;
    qtx_status st = QTX_OK;
    for (int w : wgs) {
        for (uint32_t k : kops) {
            b->tuning = qtx_batch_tuning{w, k};
            if ((st = rebuild(b)) != QTX_OK) break;
            if ((st = qtx_batch_launch(b, stream)) != QTX_OK) break;  // warm-up
            e = hipEventRecord(t0, s);
            for (int r = 0; r < repeats && st == QTX_OK && e == hipSuccess; r++) st = qtx_batch_launch(b, stream);
            if (st != QTX_OK) break;
            if (e == hipSuccess) e = hipEventRecord(t1, s);
            if (e == hipSuccess) e = hipEventSynchronize(t1);
            float ms = 0.0f;
            if (e == hipSuccess) e = hipEventElapsedTime(&ms, t0, t1);
            if (e != hipSuccess) {
                st = hip_fail(e, "autotune timing");
                break;
            }
            ms /= (float)repeats;
            if (ms < chosen_ms) {
                chosen_ms = ms;
                chosen = b->tuning;
            }
        }
        if (st != QTX_OK) break;
    }
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    if (st != QTX_OK) return st;
    b->tuning = chosen;
    if ((st = rebuild(b)) != QTX_OK) return st;
    if (best) *best = chosen;
    if (best_ms) *best_ms = chosen_ms;
    return QTX_OK;
}

qtx_status qtx_batch_run(qtx_ctx *ctx, const qtx_job *jobs, size_t njobs, void *stream) {
    qtx_batch *b = nullptr;
    qtx_status st = qtx_batch_create(ctx, jobs, njobs, &b);
    if (st != QTX_OK) return st;
    st = qtx_batch_launch(b, stream);
    if (st == QTX_OK) {
        const hipError_t e = hipStreamSynchronize((hipStream_t)stream);
        if (e != hipSuccess) st = hip_fail(e, "hipStreamSynchronize");
    }
    qtx_batch_destroy(b);
    return st;
}

qtx_status qtx_batch_run_host(qtx_ctx *ctx, const qtx_job *jobs, const size_t *dst_bytes, size_t njobs) {
    qtx_status st = check_ctx(ctx);
    if (st != QTX_OK) return st;
    if (!jobs || !dst_bytes || njobs == 0) return fail(QTX_EINVAL, "empty batch");
    // sizes first (validates formats and dimensions)
    std::vector<Prepared> prep(njobs);
    std::vector<size_t> src_off(njobs), dst_off(njobs);
    size_t total = 0;
    auto align = [](size_t n) { return (n + 255) & ~(size_t)255; };
    for (size_t i = 0; i < njobs; i++) {
        if ((st = prepare(jobs[i], i, prep[i])) != QTX_OK) return st;
        if (dst_bytes[i] < prep[i].dst_bytes)
            return fail(QTX_EINVAL, "job %zu: output buffer holds %zu bytes, %zu needed", i, dst_bytes[i],
                        prep[i].dst_bytes);
        src_off[i] = total;
        total += align(prep[i].src_bytes);
    }
    for (size_t i = 0; i < njobs; i++) {
        dst_off[i] = total;
        total += align(prep[i].dst_bytes);
    }
    std::lock_guard<std::mutex> lk(ctx->mu);
    if ((st = ensure(ctx->d_in, ctx->in_cap, total)) != QTX_OK) return st;
    std::vector<qtx_job> dev(jobs, jobs + njobs);
    // the sources lie in one range of the device buffer and the destinations in
    // another: one staged transfer each way (parts only read on the way up)
    std::vector<rt::HostPart> up(njobs), down(njobs);
    for (size_t i = 0; i < njobs; i++) {
        up[i] = {const_cast<uint8_t *>(static_cast<const uint8_t *>(jobs[i].src)), src_off[i], prep[i].src_bytes};
        down[i] = {static_cast<uint8_t *>(jobs[i].dst), dst_off[i] - dst_off[0], prep[i].dst_bytes};
        dev[i].src = ctx->d_in + src_off[i];
        dev[i].dst = ctx->d_in + dst_off[i];
    }
    QTX_HIP(ctx->stager.upload_parts(ctx->d_in, up.data(), njobs, src_off[njobs - 1] + prep[njobs - 1].src_bytes,
                                     ctx->stream));
    std::vector<BatchJob> table;  // read by the copy to the device until the stream is synchronised
    st = run_once(ctx, dev, ctx->stream, table);
    if (st == QTX_OK) {
        const hipError_t e = ctx->stager.download_parts(down.data(), njobs, ctx->d_in + dst_off[0],
                                                        dst_off[njobs - 1] - dst_off[0] + prep[njobs - 1].dst_bytes,
                                                        ctx->stream);
        if (e != hipSuccess) st = hip_fail(e, "download");
    }
    const hipError_t e = hipStreamSynchronize(ctx->stream);
    if (st == QTX_OK && e != hipSuccess) st = hip_fail(e, "hipStreamSynchronize");
    return st;
}

}  // extern "C"
