// capi.cpp -- the C ABI (include/qtx.h) over the gfx950 kernels.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iterator>
#include <mutex>
#include <string>
#include <vector>

#include "kernels.h"
#include "qtx_internal.h"
#include "runtime.h"

// Device memory of qtx_bc1_encode_device_noalloc: the lane kernels' work queue
// words (kLaneQueueWords) followed by park records for up to
// QTX_NOALLOC_MAX_BLOCKS blocks (or, for the static lane kernel, its block lists,
// at most 1 MiB there), 8 MiB + 1 KiB in all.  Each device holds its own
// copy from the library's load until the process ends; host code reaches it only
// through the address hipGetSymbolAddress gives for the current device (kept in
// qtx_ctx::d_scratch), never through this name.
namespace {
constexpr size_t kScratchQueueBytes = sizeof(uint32_t) * qtx::kLaneQueueWords;
constexpr size_t kScratchBytes = kScratchQueueBytes + qtx::kParkRecordBytes * (size_t)QTX_NOALLOC_MAX_BLOCKS;
static_assert(sizeof(uint32_t) * qtx::kLaneQueues * qtx::list_capacity(QTX_NOALLOC_MAX_BLOCKS) <=
                  qtx::kParkRecordBytes * (size_t)QTX_NOALLOC_MAX_BLOCKS,
              "the block lists fit where the park records go");
}  // namespace
static __device__ __align__(256) uint8_t scratch[kScratchBytes];

namespace qtx {
namespace rt {

namespace {
thread_local std::string g_err;
}  // namespace

const char *last_error() { return g_err.c_str(); }

qtx_status fail(qtx_status st, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return st;
}

qtx_status hip_fail(hipError_t e, const char *what) {
    return fail(QTX_EDEVICE, "%s: %s", what, hipGetErrorString(e));
}

qtx_status check_dims(int32_t w, int32_t h) {
    // RawTexture/BlockTexture constructors (quicktex/Texture.h:56-57)
    if (w <= 0) return fail(QTX_EINVAL, "Texture width must be greater than 0");
    if (h <= 0) return fail(QTX_EINVAL, "Texture height must be greater than 0");
    const long long bw = (w + 3LL) / 4, bh = (h + 3LL) / 4;
    if ((long long)w * h > (1LL << 31) || bw * bh * 16 >= (1LL << 31))
        return fail(QTX_ERANGE, "texture of %d x %d exceeds the 2^31-texel limit of this library", w, h);
    return QTX_OK;
}

size_t block_bytes(int32_t w, int32_t h, int bytes_per_block) {
    return (size_t)((w + 3) / 4) * (size_t)((h + 3) / 4) * (size_t)bytes_per_block;
}

size_t raw_bytes(int32_t w, int32_t h) { return (size_t)w * (size_t)h * 4; }

namespace {
// A stream's identity for stream-ordered reuse of queue sets (the null stream: ~0).
uint64_t stream_id(hipStream_t s) {
    if (!s) return ~0ull;
    unsigned long long id = 0;
    if (hipStreamGetId(s, &id) != hipSuccess) {
        (void)hipGetLastError();
        return 0;  // unknown: never matched
    }
    return (uint64_t)id + 1;
}
}  // namespace

QueueLease::QueueLease(qtx_ctx *ctx, hipStream_t stream, bool wanted) : ctx_(ctx), stream_(stream) {
    if (!wanted) return;
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(stream, &cap) != hipSuccess || cap != hipStreamCaptureStatusNone) return;
    ctx->queue_mu.lock();
    locked_ = true;
    sid_ = stream_id(stream);
    // a set whose last launch has completed; else one whose last launch was on this
    // stream: stream order runs that launch (and its drain, which leaves the words
    // zero) before this one, so the set is free by the time this launch starts
    int same_stream = -1;
    for (int i = 0; i < qtx_ctx::kQueueSlots; i++) {
        hipEvent_t &ev = ctx->queue_done[i];
        if (!ev && hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess) {
            ev = nullptr;
            break;
        }
        if (hipEventQuery(ev) == hipSuccess) {  // never recorded, or its launch is done
            slot_ = i;
            break;
        }
        if (same_stream < 0 && sid_ != 0 && ctx->queue_stream[i] == sid_) same_stream = i;
    }
    (void)hipGetLastError();  // hipEventQuery's not-ready status
    if (slot_ < 0) slot_ = same_stream;
    if (slot_ >= 0) words = ctx->d_queues + (size_t)kLaneQueueWords * slot_;
}

void QueueLease::reserve_park(size_t bytes) {
    park = nullptr;
    if (slot_ < 0 || bytes == 0) return;
    void *&buf = ctx_->d_park[slot_];
    size_t &cap = ctx_->park_cap[slot_];
    if (cap < bytes) {
        // stream-ordered: the old buffer is released after the work queued on this
        // stream so far (the set's last launch completed, or ran on this stream), and
        // neither call waits for the device
        if (buf) (void)hipFreeAsync(buf, stream_);
        buf = nullptr;
        cap = 0;
        if (hipMallocAsync(&buf, bytes, stream_) != hipSuccess) {
            buf = nullptr;
            (void)hipGetLastError();
            return;
        }
        cap = bytes;
    }
    park = buf;
}

hipError_t QueueLease::commit() {
    hipError_t e = hipSuccess;
    if (slot_ >= 0) {
        e = hipEventRecord(ctx_->queue_done[slot_], stream_);
        ctx_->queue_stream[slot_] = sid_;
    }
    slot_ = -1;
    if (locked_) ctx_->queue_mu.unlock();
    locked_ = false;
    return e;
}

QueueLease::~QueueLease() { (void)commit(); }

qtx_status check_ctx(qtx_ctx *ctx) {
    if (!ctx) return fail(QTX_EINVAL, "null context");
    QTX_HIP(hipSetDevice(ctx->device));
    return QTX_OK;
}

qtx_status to_settings(const qtx_bc1_params *p, BC1Settings &s, int &interp) {
    if (!p) return fail(QTX_EINVAL, "null params");
    qtx_bc1_params c = *p;
    qtx_status st = qtx_bc1_params_clamp(&c);
    if (st != QTX_OK) return st;
    s.color_mode = c.color_mode;
    s.error_mode = c.error_mode;
    s.endpoint_mode = c.endpoint_mode;
    s.ls_passes = c.two_ls_passes ? 2 : 1;
    // orderings are >= 1, so every block needs its error and two_ep applies
    // (BC1Encoder.cpp:269-282)
    s.ep_passes = c.two_ep_passes ? 2 : 1;
    s.cf_passes = c.two_cf_passes ? 2 : 1;
    s.exhaustive = c.exhaustive ? 1 : 0;
    s.power_iterations = c.power_iterations;
    s.search_rounds = c.search_rounds;
    s.orderings4 = c.orderings4;
    s.orderings3 = c.orderings3;
    interp = c.interpolator;
    return QTX_OK;
}

qtx_status ensure(uint8_t *&buf, size_t &cap, size_t need) {
    if (need <= cap) return QTX_OK;
    if (buf) (void)hipFree(buf);
    buf = nullptr;
    cap = 0;
    const size_t n = std::max(need, (size_t)1 << 20);
    hipError_t e = hipMalloc(&buf, n);
    if (e != hipSuccess) return fail(QTX_ENOMEM, "hipMalloc(%zu): %s", n, hipGetErrorString(e));
    cap = n;
    return QTX_OK;
}

}  // namespace rt
}  // namespace qtx

using namespace qtx::rt;

namespace {

// The band resources of a context (qtx_ctx::copy_stream and its events).
void release_bands(qtx_ctx *ctx) {
    if (ctx->copy_stream) {
        (void)hipStreamSynchronize(ctx->copy_stream);
        (void)hipStreamDestroy(ctx->copy_stream);
        ctx->copy_stream = nullptr;
    }
    for (hipEvent_t *evs : {ctx->band_up, ctx->band_done})
        for (int k = 0; k < qtx_ctx::kMaxBands; k++)
            if (evs[k]) (void)hipEventDestroy(evs[k]), evs[k] = nullptr;
}

// Makes them, all or nothing; the device is current.
hipError_t setup_bands(qtx_ctx *ctx) {
    if (ctx->copy_stream) return hipSuccess;
    hipError_t e = hipSuccess;
    for (int k = 0; k < qtx_ctx::kMaxBands && e == hipSuccess; k++) {
        e = hipEventCreateWithFlags(&ctx->band_up[k], hipEventDisableTiming);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&ctx->band_done[k], hipEventDisableTiming);
    }
    // the stream last: its being set means the events are there
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&ctx->copy_stream, hipStreamNonBlocking);
    if (e != hipSuccess) release_bands(ctx);
    return e;
}

}  // namespace

namespace qtx {
namespace rt {
hipError_t ensure_copy_stream(qtx_ctx *ctx) { return setup_bands(ctx); }
}  // namespace rt
}  // namespace qtx

extern "C" {

const char *qtx_last_error(void) { return last_error(); }

qtx_status qtx_bc1_params_clamp(qtx_bc1_params *p) {
    if (!p) return fail(QTX_EINVAL, "null params");
    if (p->interpolator < 0 || p->interpolator > 3) return fail(QTX_EINVAL, "Invalid interpolator type");
    if (p->color_mode != QTX_COLOR_FOUR && p->color_mode != QTX_COLOR_THREE && p->color_mode != QTX_COLOR_THREE_BLACK)
        return fail(QTX_EINVAL, "Encoder color mode must be FourColor, ThreeColor, or ThreeColorBlack");
    if (p->error_mode < 0 || p->error_mode > 3) return fail(QTX_EINVAL, "invalid error mode %d", p->error_mode);
    if (p->endpoint_mode < 0 || p->endpoint_mode > 3) return fail(QTX_EINVAL, "invalid endpoint mode %d", p->endpoint_mode);
    // BC1Encoder.cpp:251-258
    p->orderings4 = std::min(std::max(p->orderings4, 1u), (uint32_t)qtx::kBestOrders4);
    p->orderings3 = std::min(std::max(p->orderings3, 1u), (uint32_t)qtx::kBestOrders3);
    p->power_iterations = std::min(std::max(p->power_iterations, 4u), 10u);
    return QTX_OK;
}

qtx_status qtx_bc1_params_init(qtx_bc1_params *p, uint32_t level, int32_t color_mode, int32_t interpolator) {
    if (!p) return fail(QTX_EINVAL, "null params");
    if (color_mode != QTX_COLOR_FOUR && color_mode != QTX_COLOR_THREE && color_mode != QTX_COLOR_THREE_BLACK)
        return fail(QTX_EINVAL, "Encoder color mode must be FourColor, ThreeColor, or ThreeColorBlack");
    if (interpolator < 0 || interpolator > 3) return fail(QTX_EINVAL, "Invalid interpolator type");
    if (level > 19) return fail(QTX_EINVAL, "Level out of range, bust be between 0 and 18 inclusive");
    // Presets of BC1Encoder::SetLevel (quicktex/s3tc/bc1/BC1Encoder.cpp:74-249):
    //   level: endpoint, error, power iterations, ls2, ep2, cf2, exhaustive, o4, o3, search rounds
    struct Preset {
        int8_t ep, err, pi, ls2, ep2, cf2, ex;
        uint8_t o4, o3;
        uint16_t search;
    };
    static const Preset presets[20] = {
        {QTX_EP_BOUNDING_BOX_INT, QTX_ERROR_CHECK2, 4, 0, 0, 0, 0, 0, 0, 0},
        {QTX_EP_LEAST_SQUARES, QTX_ERROR_CHECK2, 4, 0, 0, 0, 0, 0, 0, 0},
        {QTX_EP_PCA, QTX_ERROR_CHECK2, 4, 0, 0, 0, 0, 0, 0, 0},
        {QTX_EP_PCA, QTX_ERROR_CHECK2, 4, 1, 0, 0, 0, 0, 0, 0},
        {QTX_EP_PCA, QTX_ERROR_FULL, 6, 1, 0, 0, 0, 0, 0, 0},
        {QTX_EP_PCA, QTX_ERROR_FASTER, 4, 1, 0, 0, 0, 0, 0, 0},
        {QTX_EP_PCA, QTX_ERROR_FASTER, 4, 1, 0, 0, 0, 1, 1, 0},
        {QTX_EP_PCA, QTX_ERROR_FASTER, 4, 1, 0, 0, 0, 4, 1, 0},
        {QTX_EP_PCA, QTX_ERROR_FASTER, 4, 1, 0, 0, 0, 8, 1, 0},
        {QTX_EP_PCA, QTX_ERROR_CHECK2, 4, 1, 0, 0, 0, 11, 3, 0},
        {QTX_EP_PCA, QTX_ERROR_CHECK2, 4, 1, 0, 0, 0, 20, 8, 0},
        {QTX_EP_PCA, QTX_ERROR_CHECK2, 4, 1, 0, 0, 0, 28, 16, 0},
        {QTX_EP_PCA, QTX_ERROR_CHECK2, 4, 1, 0, 0, 0, 32, 32, 0},
        {QTX_EP_PCA, QTX_ERROR_FULL, 6, 1, 1, 0, 0, 32, 32, 20},
        {QTX_EP_PCA, QTX_ERROR_FULL, 6, 1, 1, 0, 0, 32, 32, 32},
        {QTX_EP_PCA, QTX_ERROR_FULL, 6, 1, 1, 0, 0, 56, 32, 32},
        {QTX_EP_PCA, QTX_ERROR_FULL, 6, 1, 1, 0, 0, 80, 32, 256},
        {QTX_EP_PCA, QTX_ERROR_FULL, 4, 1, 1, 0, 0, 128, 32, 256},  // level 17 leaves power iterations at 4
        {QTX_EP_PCA, QTX_ERROR_FULL, 6, 1, 1, 1, 0, 128, 32, 256},
        {QTX_EP_PCA, QTX_ERROR_FULL, 6, 1, 1, 1, 1, 128, 32, 256},
    };
    const Preset &s = presets[level];
    p->interpolator = interpolator;
    p->color_mode = color_mode;
    p->error_mode = s.err;
    p->endpoint_mode = s.ep;
    p->two_ls_passes = s.ls2;
    p->two_ep_passes = s.ep2;
    p->two_cf_passes = s.cf2;
    p->exhaustive = s.ex;
    p->power_iterations = (uint32_t)s.pi;
    p->search_rounds = s.search;
    p->orderings4 = s.o4;
    p->orderings3 = s.o3;
    return qtx_bc1_params_clamp(p);
}

qtx_status qtx_ctx_create(int32_t device, qtx_ctx **out) {
    if (!out) return fail(QTX_EINVAL, "null output pointer");
    *out = nullptr;
    int n = 0;
    QTX_HIP(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) return fail(QTX_EINVAL, "device %d out of range (%d devices)", device, n);
    QTX_HIP(hipSetDevice(device));
    auto *ctx = new qtx_ctx;
    ctx->device = device;
    hipError_t e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&ctx->cu_count, hipDeviceAttributeMultiprocessorCount, device);
    if (e == hipSuccess) e = hipMalloc(&ctx->d_tables, sizeof(qtx::DeviceTables));
    if (e == hipSuccess)
        e = hipMemcpy(ctx->d_tables, &qtx::host_tables(), sizeof(qtx::DeviceTables), hipMemcpyHostToDevice);
    const size_t qbytes = sizeof(uint32_t) * qtx::kLaneQueueWords * qtx_ctx::kQueueSlots;
    if (e == hipSuccess) e = hipMalloc(&ctx->d_queues, qbytes);
    if (e == hipSuccess) e = hipMemset(ctx->d_queues, 0, qbytes);
    // this device's copy of the static scratch array (the device is current)
    if (e == hipSuccess) e = hipGetSymbolAddress((void **)&ctx->d_scratch, HIP_SYMBOL(scratch));
    if (e == hipSuccess && !ctx->d_scratch) e = hipErrorInvalidSymbol;
    if (e != hipSuccess) {
        qtx_ctx_destroy(ctx);
        return hip_fail(e, "context set-up");
    }
    *out = ctx;
    return QTX_OK;
}

qtx_status qtx_ctx_destroy(qtx_ctx *ctx) {
    if (!ctx) return QTX_OK;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    if (ctx->d_in) (void)hipFree(ctx->d_in);
    if (ctx->d_out) (void)hipFree(ctx->d_out);
    if (ctx->d_flag) (void)hipFree(ctx->d_flag);
    if (ctx->d_once) (void)hipFree(ctx->d_once);
    if (ctx->d_tables) (void)hipFree(ctx->d_tables);
    for (hipEvent_t &ev : ctx->queue_done) {
        if (!ev) continue;
        (void)hipEventSynchronize(ev);  // the last launch that used the slot
        (void)hipEventDestroy(ev);
    }
    if (ctx->d_queues) (void)hipFree(ctx->d_queues);
    for (void *p : ctx->d_park)
        if (p) (void)hipFreeAsync(p, ctx->stream);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    release_bands(ctx);
    delete ctx;
    return QTX_OK;
}

qtx_status qtx_ctx_set_bc1_encoder(qtx_ctx *ctx, int32_t which) {
    if (!ctx) return fail(QTX_EINVAL, "null context");
    if (which != QTX_BC1_AUTO && which != QTX_BC1_LANE && which != QTX_BC1_GROUP)
        return fail(QTX_EINVAL, "unknown BC1 encoder %d", which);
    ctx->bc1_path = which;
    return QTX_OK;
}

qtx_status qtx_ctx_synchronize(qtx_ctx *ctx) {
    qtx_status st = check_ctx(ctx);
    if (st != QTX_OK) return st;
    QTX_HIP(hipStreamSynchronize(ctx->stream));
    return QTX_OK;
}

// ---------------------------------------------------------------------------
// device-resident entry points

qtx_status qtx_bc1_encode_device(qtx_ctx *ctx, const qtx_bc1_params *params, const uint8_t *d_rgba, int32_t width,
                                 int32_t height, uint8_t *d_blocks, void *stream) {
    qtx_status st = check_ctx(ctx);
    if (st == QTX_OK) st = check_dims(width, height);
    if (st != QTX_OK) return st;
    qtx::BC1Settings s;
    int interp = 0;
    if ((st = to_settings(params, s, interp)) != QTX_OK) return st;
    if (!d_rgba || !d_blocks) return fail(QTX_EINVAL, "null buffer");
    QueueLease lease(ctx, (hipStream_t)stream, qtx::wants_lane_queue(s, width, height, ctx->bc1_path));
    lease.reserve_park(qtx::bc1_park_bytes(s, width, height, ctx->bc1_path));
    QTX_HIP(qtx::launch_bc1_encode(d_rgba, width, height, d_blocks, ctx->d_tables, s, interp, false, ctx->bc1_path,
                                   lease.words, lease.park, ctx->cu_count, (hipStream_t)stream));
    QTX_HIP(lease.commit());
    return QTX_OK;
}

qtx_status qtx_bc1_encode_device_noalloc(qtx_ctx *ctx, const qtx_bc1_params *params, const uint8_t *d_rgba,
                                         int32_t width, int32_t height, uint8_t *d_blocks, void *stream) {
    qtx_status st = check_ctx(ctx);
    if (st == QTX_OK) st = check_dims(width, height);
    if (st != QTX_OK) return st;
    const uint64_t nblocks = (uint64_t)((width + 3) / 4) * (uint64_t)((height + 3) / 4);
    if (nblocks > (uint64_t)QTX_NOALLOC_MAX_BLOCKS)
        return fail(QTX_ERANGE, "texture of %d x %d has %llu blocks; the no-allocation encoder takes at most %d",
                    width, height, (unsigned long long)nblocks, QTX_NOALLOC_MAX_BLOCKS);
    qtx::BC1Settings s;
    int interp = 0;
    if ((st = to_settings(params, s, interp)) != QTX_OK) return st;
    if (!d_rgba || !d_blocks) return fail(QTX_EINVAL, "null buffer");
    if (!ctx->d_scratch) return fail(QTX_EDEVICE, "no scratch memory on device %d", ctx->device);
    const hipStream_t hs = (hipStream_t)stream;
    uint32_t *words = nullptr;
    void *park = nullptr;
    if (qtx::wants_lane_queue(s, width, height, ctx->bc1_path)) {
        // the queue and park counters start at zero (the previous call's kernels
        // leave them so, but a call cut short may not have)
        words = reinterpret_cast<uint32_t *>(ctx->d_scratch);
        QTX_HIP(qtx::launch_zero_words(words, (uint32_t)qtx::kLaneQueueWords, hs));
        if (qtx::bc1_park_bytes(s, width, height, ctx->bc1_path) > 0) park = ctx->d_scratch + kScratchQueueBytes;
    }
    QTX_HIP(qtx::launch_bc1_encode(d_rgba, width, height, d_blocks, ctx->d_tables, s, interp, false, ctx->bc1_path,
                                   words, park, ctx->cu_count, hs));
    return QTX_OK;
}

qtx_status qtx_bc3_encode_device(qtx_ctx *ctx, const qtx_bc1_params *params, const uint8_t *d_rgba, int32_t width,
                                 int32_t height, uint8_t *d_blocks, void *stream) {
    qtx_status st = check_ctx(ctx);
    if (st == QTX_OK) st = check_dims(width, height);
    if (st != QTX_OK) return st;
    qtx::BC1Settings s;
    int interp = 0;
    if ((st = to_settings(params, s, interp)) != QTX_OK) return st;
    if (s.color_mode != QTX_COLOR_FOUR) return fail(QTX_EINVAL, "BC3 colour blocks are always FourColor");
    if (!d_rgba || !d_blocks) return fail(QTX_EINVAL, "null buffer");
    QueueLease lease(ctx, (hipStream_t)stream, qtx::wants_lane_queue(s, width, height, ctx->bc1_path));
    lease.reserve_park(qtx::bc1_park_bytes(s, width, height, ctx->bc1_path));
    QTX_HIP(qtx::launch_bc1_encode(d_rgba, width, height, d_blocks, ctx->d_tables, s, interp, true, ctx->bc1_path,
                                   lease.words, lease.park, ctx->cu_count, (hipStream_t)stream));
    QTX_HIP(lease.commit());
    return QTX_OK;
}

qtx_status qtx_bc4_encode_device(qtx_ctx *ctx, uint32_t channel, const uint8_t *d_rgba, int32_t width, int32_t height,
                                 uint8_t *d_blocks, void *stream) {
    qtx_status st = check_ctx(ctx);
    if (st == QTX_OK) st = check_dims(width, height);
    if (st != QTX_OK) return st;
    if (channel >= 4) return fail(QTX_EINVAL, "Channel out of range");
    if (!d_rgba || !d_blocks) return fail(QTX_EINVAL, "null buffer");
    QTX_HIP(qtx::launch_bc4_encode(d_rgba, width, height, d_blocks, (int)channel, 0, false, (hipStream_t)stream));
    return QTX_OK;
}

qtx_status qtx_bc5_encode_device(qtx_ctx *ctx, uint32_t chan0, uint32_t chan1, const uint8_t *d_rgba, int32_t width,
                                 int32_t height, uint8_t *d_blocks, void *stream) {
    qtx_status st = check_ctx(ctx);
    if (st == QTX_OK) st = check_dims(width, height);
    if (st != QTX_OK) return st;
    if (chan0 >= 4 || chan1 >= 4) return fail(QTX_EINVAL, "Channel out of range");
    if (!d_rgba || !d_blocks) return fail(QTX_EINVAL, "null buffer");
    QTX_HIP(qtx::launch_bc4_encode(d_rgba, width, height, d_blocks, (int)chan0, (int)chan1, true, (hipStream_t)stream));
    return QTX_OK;
}

static qtx_status decode_device(qtx_ctx *ctx, int format, int interp, int p0, int p1, const uint8_t *d_blocks,
                                int32_t width, int32_t height, uint8_t *d_rgba, void *stream) {
    qtx_status st = check_ctx(ctx);
    if (st == QTX_OK) st = check_dims(width, height);
    if (st != QTX_OK) return st;
    if (interp < 0 || interp > 3) return fail(QTX_EINVAL, "Invalid interpolator type");
    if (!d_rgba || !d_blocks) return fail(QTX_EINVAL, "null buffer");
    QTX_HIP(qtx::launch_decode(format, d_blocks, width, height, d_rgba, interp, p0, p1, (hipStream_t)stream));
    return QTX_OK;
}

qtx_status qtx_bc1_decode_device(qtx_ctx *ctx, int32_t interpolator, int32_t write_alpha, const uint8_t *d_blocks,
                                 int32_t width, int32_t height, uint8_t *d_rgba, void *stream) {
    return decode_device(ctx, QTX_FMT_BC1, interpolator, write_alpha != 0, 0, d_blocks, width, height, d_rgba, stream);
}

qtx_status qtx_bc3_decode_device(qtx_ctx *ctx, int32_t interpolator, const uint8_t *d_blocks, int32_t width,
                                 int32_t height, uint8_t *d_rgba, void *stream) {
    return decode_device(ctx, QTX_FMT_BC3, interpolator, 0, 0, d_blocks, width, height, d_rgba, stream);
}

qtx_status qtx_bc4_decode_device(qtx_ctx *ctx, uint32_t channel, const uint8_t *d_blocks, int32_t width,
                                 int32_t height, uint8_t *d_rgba, void *stream) {
    if (channel >= 4) return fail(QTX_EINVAL, "Channel out of range");
    return decode_device(ctx, QTX_FMT_BC4, 0, (int)channel, 0, d_blocks, width, height, d_rgba, stream);
}

qtx_status qtx_bc5_decode_device(qtx_ctx *ctx, uint32_t chan0, uint32_t chan1, const uint8_t *d_blocks, int32_t width,
                                 int32_t height, uint8_t *d_rgba, void *stream) {
    if (chan0 >= 4 || chan1 >= 4) return fail(QTX_EINVAL, "Channel out of range");
    return decode_device(ctx, QTX_FMT_BC5, 0, (int)chan0, (int)chan1, d_blocks, width, height, d_rgba, stream);
}

// ---------------------------------------------------------------------------
// host entry points: stage through the context's device buffers (large transfers
// through its pinned chunks, staging.h)

}  // extern "C"

namespace {

// Host calls on pageable buffers whose input and output together take at most
// this many bytes skip the DMA engine: the input is copied into the context's
// pinned chunks, the kernels read it and write their output there over the link,
// and the output is copied out.  A transfer by the DMA engine costs 8-15 us of set-up and wait
// however small (measured: 256 KiB up 12-15 us, 32 KiB down 8-9 us, a launch and
// its wait 11 us), which is most of a small call: host calls of 128x128 took
// 19-50 us this way against 60-93 us through the DMA engine, 256x256 25-61 against
// 40-84, 512x512 71-133 against 82-159; at 724x724 the two are even, and from
// 1024x1024 on the kernels' reads and writes over the link make decodes slower.
// QTX_HOST_DIRECT_MAX (bytes) sets another limit; 0 sends every call through the
// DMA engine.
constexpr size_t kDirectMaxBytes = (size_t)3 << 19;  // 1.5 MiB: BC1 up to 590x590

size_t direct_max_bytes() {
    static const size_t v = [] {
        if (const char *e = std::getenv("QTX_HOST_DIRECT_MAX")) {
            char *end;
            const unsigned long long n = std::strtoull(e, &end, 10);
            if (end != e) return std::min((size_t)n, qtx::rt::Stager::kSlots * qtx::rt::Stager::kChunk);
        }
        return kDirectMaxBytes;
    }();
    return v;
}

template <class Launch>
qtx_status host_call(qtx_ctx *ctx, const uint8_t *in, size_t in_bytes, uint8_t *out, size_t out_need, size_t out_bytes,
                     Launch &&launch) {
    if (!ctx) return fail(QTX_EINVAL, "null context");
    if (!in || !out) return fail(QTX_EINVAL, "null buffer");
    if (out_bytes < out_need) return fail(QTX_EINVAL, "output buffer holds %zu bytes, %zu needed", out_bytes, out_need);
    std::lock_guard<std::mutex> lk(ctx->mu);
    QTX_HIP(hipSetDevice(ctx->device));
    const size_t in_span = (in_bytes + 255) & ~(size_t)255;  // the output starts 256-byte aligned
    const bool small = in_span + out_need <= direct_max_bytes();
    const bool in_pinned = small && qtx::rt::host_pinned(in), out_pinned = small && qtx::rt::host_pinned(out);
    if (in_pinned && out_pinned) {  // the kernels read and write the caller's pinned buffers
        const qtx_status st = launch(in, out, (void *)ctx->stream);
        const hipError_t e = hipStreamSynchronize(ctx->stream);
        if (st != QTX_OK) return st;
        if (e != hipSuccess) return hip_fail(e, "hipStreamSynchronize");
        return QTX_OK;
    }
    uint8_t *pin = nullptr;
    // both buffers pageable: a pinned one beside a pageable one goes to the DMA engine
    // directly, which beats a copy through the chunks (the served mix from pinned
    // buffers: small decodes 0.06 ms by DMA against 0.09-0.12 ms through the chunks,
    // and the CPU's first write into a caller's untouched pinned output took 0.4-0.5
    // ms more)
    if (small && !in_pinned && !out_pinned && ctx->stager.direct(in_span + out_need, &pin) == hipSuccess) {
        std::memcpy(pin, in, in_bytes);
        const qtx_status st = launch(pin, pin + in_span, (void *)ctx->stream);
        const hipError_t e = hipStreamSynchronize(ctx->stream);
        if (st != QTX_OK) return st;
        if (e != hipSuccess) return hip_fail(e, "hipStreamSynchronize");
        std::memcpy(out, pin + in_span, out_need);
        return QTX_OK;
    }
    qtx_status st = ensure(ctx->d_in, ctx->in_cap, in_bytes);
    if (st == QTX_OK) st = ensure(ctx->d_out, ctx->out_cap, out_need);
    if (st != QTX_OK) return st;
    QTX_HIP(ctx->stager.upload(ctx->d_in, in, in_bytes, ctx->stream));
    st = launch(ctx->d_in, ctx->d_out, (void *)ctx->stream);
    if (st != QTX_OK) {
        (void)hipStreamSynchronize(ctx->stream);
        return st;
    }
    QTX_HIP(ctx->stager.download(out, ctx->d_out, out_need, ctx->stream));
    QTX_HIP(hipStreamSynchronize(ctx->stream));
    return QTX_OK;
}

// Host encodes of kBandMinBytes of texels and more are cut into bands of block
// rows.  Band k is uploaded on the copy stream and encoded on ctx->stream once its
// rows have landed, so the upload of the next bands overlaps this band's encode.
// Blocks are encoded independently of each other
// from their own texels (a ragged last block row's padding aside, handled in
// host_encode), so the bands' blocks are those of one launch over the whole
// texture.  The blocks are downloaded band by band once every band is enqueued.
constexpr size_t kBandMinBytes = (size_t)16 << 20;
// A band's blocks are staged from this size (the runtime's path for pageable
// memory runs at about half the rate)
constexpr size_t kBandStageBytes = (size_t)1 << 20;
// Band sizes in sixteenths of the block rows: the first quarter starts the GPU
// while the rest is uploaded.  Measured from pageable memory on BC1 4096x4096
// against one band: level 18 4.64 -> 3.90 ms, level 10 2.47 -> 2.30 ms, level 5
// 2.07 -> 2.01 ms (2048x2048: level 18 1.56 -> 1.58 ms, level 10 0.74 -> 0.71 ms).
// More bands leave more launch tails, which one stream does not overlap.
// QTX_HOST_BANDS ("2,6,8", ...) sets another plan; "16" is one band.
constexpr int kBandPlan[] = {4, 12};

struct BandPlan {
    int n = 0;
    int parts[qtx_ctx::kMaxBands] = {};
};

BandPlan read_band_plan() {
    BandPlan p;
    if (const char *e = std::getenv("QTX_HOST_BANDS")) {
        int sum = 0;
        for (const char *c = e; p.n < qtx_ctx::kMaxBands - 1;) {  // one more unit for a ragged last block row
            char *end;
            const long v = std::strtol(c, &end, 10);
            if (end == c || v <= 0 || v > 16) break;
            p.parts[p.n++] = (int)v;
            sum += (int)v;
            if (*end != ',') break;
            c = end + 1;
        }
        if (sum == 16) return p;
    }
    p.n = (int)std::size(kBandPlan);
    std::copy(std::begin(kBandPlan), std::end(kBandPlan), p.parts);
    return p;
}

template <class BandLaunch>
qtx_status host_encode(qtx_ctx *ctx, const uint8_t *rgba, int32_t w, int32_t h, int block_size, uint8_t *out,
                       size_t out_bytes, BandLaunch &&launch) {
    const size_t in_bytes = raw_bytes(w, h), out_need = block_bytes(w, h, block_size);
    const int bh = (h + 3) / 4, full = h / 4;  // block rows; those with four texel rows
    if (in_bytes < kBandMinBytes || full < 2)
        return host_call(ctx, rgba, in_bytes, out, out_need, out_bytes,
                         [&](const uint8_t *di, uint8_t *dout, void *s) { return launch(di, h, dout, s); });
    if (!ctx) return fail(QTX_EINVAL, "null context");
    if (!rgba || !out) return fail(QTX_EINVAL, "null buffer");
    if (out_bytes < out_need) return fail(QTX_EINVAL, "output buffer holds %zu bytes, %zu needed", out_bytes, out_need);
    std::lock_guard<std::mutex> lk(ctx->mu);
    QTX_HIP(hipSetDevice(ctx->device));
    const size_t row_in = (size_t)w * 16, row_out = (size_t)((w + 3) / 4) * block_size;
    qtx_status st = ensure(ctx->d_in, ctx->in_cap, bh * row_in);
    if (st == QTX_OK) st = ensure(ctx->d_out, ctx->out_cap, out_need);
    if (st != QTX_OK) return st;
    QTX_HIP(setup_bands(ctx));

    // Units of work: bands of the block rows with four texel rows, then, for a
    // height that is not a multiple of 4, the last block row on its own.  The
    // reference pads that row by wrapping to the texture's top rows (texel row y
    // reads row y % h, quicktex/Texture.h:107), so its four rows are put together
    // here, in that order, and encoded as a texture of height 4.
    struct Unit {
        const uint8_t *src;
        size_t in_off, in_len, out_off, out_len;
        int32_t rows;
    };
    Unit units[qtx_ctx::kMaxBands];
    int n = 0;
    static const BandPlan plan = read_band_plan();
    for (int i = 0, sum = 0, r0 = 0; i < plan.n; i++) {
        sum += plan.parts[i];
        const int r1 = (int)((long long)full * sum / 16);
        if (r1 > r0) units[n++] = Unit{rgba + r0 * row_in, r0 * row_in, (r1 - r0) * row_in, r0 * row_out,
                                       (r1 - r0) * row_out, 4 * (r1 - r0)};
        r0 = std::max(r0, r1);
    }
    std::vector<uint8_t> last;
    if (full < bh) {
        last.resize(row_in);
        const size_t texel_row = (size_t)w * 4;
        for (int j = 0; j < 4; j++) {
            const int y = 4 * full + j;
            std::memcpy(last.data() + j * texel_row, rgba + (size_t)(y < h ? y : y - h) * texel_row, texel_row);
        }
        units[n++] = Unit{last.data(), full * row_in, row_in, full * row_out, row_out, 4};
    }

    hipError_t e = hipSuccess;
    for (int k = 0; k < n && e == hipSuccess && st == QTX_OK; k++) {
        const Unit &u = units[k];
        e = ctx->stager.upload(ctx->d_in + u.in_off, u.src, u.in_len, ctx->copy_stream);
        if (e == hipSuccess) e = hipEventRecord(ctx->band_up[k], ctx->copy_stream);
        if (e == hipSuccess) e = hipStreamWaitEvent(ctx->stream, ctx->band_up[k], 0);
        if (e != hipSuccess) break;
        st = launch(ctx->d_in + u.in_off, u.rows, ctx->d_out + u.out_off, (void *)ctx->stream);
        if (st == QTX_OK) e = hipEventRecord(ctx->band_done[k], ctx->stream);
    }
    for (int k = 0; k < n && e == hipSuccess && st == QTX_OK; k++) {
        const Unit &u = units[k];
        e = hipStreamWaitEvent(ctx->copy_stream, ctx->band_done[k], 0);
        if (e == hipSuccess)
            e = ctx->stager.download(out + u.out_off, ctx->d_out + u.out_off, u.out_len, ctx->copy_stream,
                                     kBandStageBytes);
    }
    // the copy stream's last work waits for every band's encode (or, after an
    // error, whatever was enqueued ends here); `last` is read until then
    const hipError_t e_copy = hipStreamSynchronize(ctx->copy_stream);
    const hipError_t e_run = hipStreamSynchronize(ctx->stream);
    if (st != QTX_OK) return st;
    if (e == hipSuccess) e = e_copy != hipSuccess ? e_copy : e_run;
    if (e != hipSuccess) return hip_fail(e, "banded host encode");
    return QTX_OK;
}

}  // namespace

extern "C" {

qtx_status qtx_bc1_encode(qtx_ctx *ctx, const qtx_bc1_params *params, const uint8_t *rgba, int32_t width,
                          int32_t height, uint8_t *out, size_t out_bytes) {
    qtx_status st = check_dims(width, height);
    if (st != QTX_OK) return st;
    return host_encode(ctx, rgba, width, height, 8, out, out_bytes,
                       [&](const uint8_t *di, int32_t band_h, uint8_t *dout, void *s) {
                           return qtx_bc1_encode_device(ctx, params, di, width, band_h, dout, s);
                       });
}

qtx_status qtx_bc3_encode(qtx_ctx *ctx, const qtx_bc1_params *params, const uint8_t *rgba, int32_t width,
                          int32_t height, uint8_t *out, size_t out_bytes) {
    qtx_status st = check_dims(width, height);
    if (st != QTX_OK) return st;
    return host_encode(ctx, rgba, width, height, 16, out, out_bytes,
                       [&](const uint8_t *di, int32_t band_h, uint8_t *dout, void *s) {
                           return qtx_bc3_encode_device(ctx, params, di, width, band_h, dout, s);
                       });
}

qtx_status qtx_bc4_encode(qtx_ctx *ctx, uint32_t channel, const uint8_t *rgba, int32_t width, int32_t height,
                          uint8_t *out, size_t out_bytes) {
    qtx_status st = check_dims(width, height);
    if (st != QTX_OK) return st;
    if (channel >= 4) return fail(QTX_EINVAL, "Channel out of range");
    return host_call(ctx, rgba, raw_bytes(width, height), out, block_bytes(width, height, 8), out_bytes,
                     [&](const uint8_t *di, uint8_t *dout, void *s) {
                         return qtx_bc4_encode_device(ctx, channel, di, width, height, dout, s);
                     });
}

qtx_status qtx_bc5_encode(qtx_ctx *ctx, uint32_t chan0, uint32_t chan1, const uint8_t *rgba, int32_t width,
                          int32_t height, uint8_t *out, size_t out_bytes) {
    qtx_status st = check_dims(width, height);
    if (st != QTX_OK) return st;
    if (chan0 >= 4 || chan1 >= 4) return fail(QTX_EINVAL, "Channel out of range");
    return host_call(ctx, rgba, raw_bytes(width, height), out, block_bytes(width, height, 16), out_bytes,
                     [&](const uint8_t *di, uint8_t *dout, void *s) {
                         return qtx_bc5_encode_device(ctx, chan0, chan1, di, width, height, dout, s);
                     });
}

qtx_status qtx_bc1_decode(qtx_ctx *ctx, int32_t interpolator, int32_t write_alpha, const uint8_t *blocks,
                          int32_t width, int32_t height, uint8_t *rgba_out, size_t out_bytes) {
    qtx_status st = check_dims(width, height);
    if (st != QTX_OK) return st;
    return host_call(ctx, blocks, block_bytes(width, height, 8), rgba_out, raw_bytes(width, height), out_bytes,
                     [&](const uint8_t *di, uint8_t *dout, void *s) {
                         return qtx_bc1_decode_device(ctx, interpolator, write_alpha, di, width, height, dout, s);
                     });
}

qtx_status qtx_bc3_decode(qtx_ctx *ctx, int32_t interpolator, const uint8_t *blocks, int32_t width, int32_t height,
                          uint8_t *rgba_out, size_t out_bytes) {
    qtx_status st = check_dims(width, height);
    if (st != QTX_OK) return st;
    return host_call(ctx, blocks, block_bytes(width, height, 16), rgba_out, raw_bytes(width, height), out_bytes,
                     [&](const uint8_t *di, uint8_t *dout, void *s) {
                         return qtx_bc3_decode_device(ctx, interpolator, di, width, height, dout, s);
                     });
}

qtx_status qtx_bc4_decode(qtx_ctx *ctx, uint32_t channel, const uint8_t *blocks, int32_t width, int32_t height,
                          uint8_t *rgba_out, size_t out_bytes) {
    qtx_status st = check_dims(width, height);
    if (st != QTX_OK) return st;
    if (channel >= 4) return fail(QTX_EINVAL, "Channel out of range");
    return host_call(ctx, blocks, block_bytes(width, height, 8), rgba_out, raw_bytes(width, height), out_bytes,
                     [&](const uint8_t *di, uint8_t *dout, void *s) {
                         return qtx_bc4_decode_device(ctx, channel, di, width, height, dout, s);
                     });
}

qtx_status qtx_bc5_decode(qtx_ctx *ctx, uint32_t chan0, uint32_t chan1, const uint8_t *blocks, int32_t width,
                          int32_t height, uint8_t *rgba_out, size_t out_bytes) {
    qtx_status st = check_dims(width, height);
    if (st != QTX_OK) return st;
    if (chan0 >= 4 || chan1 >= 4) return fail(QTX_EINVAL, "Channel out of range");
    return host_call(ctx, blocks, block_bytes(width, height, 16), rgba_out, raw_bytes(width, height), out_bytes,
                     [&](const uint8_t *di, uint8_t *dout, void *s) {
                         return qtx_bc5_decode_device(ctx, chan0, chan1, di, width, height, dout, s);
                     });
}

}  // extern "C"
