// mips.cpp -- the C ABI of the mip chain (include/qtx.h, qtx_mip*): Pillow-exact
// levels resampled on the GPU (resample.h) and, for the DDS encoder, fed to one
// batched encode without leaving the device.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>
#include <vector>

#include "batch_internal.h"
#include "resample.h"
#include "runtime.h"

using namespace qtx::rt;

namespace {

// Validates a chain request; fills the plan and its level count.
qtx_status chain_plan(int32_t w, int32_t h, int32_t levels, qtx::MipPlan &p, size_t &ws_bytes) {
    qtx_status st = check_dims(w, h);
    if (st != QTX_OK) return st;
    if (levels < 1) return fail(QTX_EINVAL, "a mip chain has at least one level (got %d)", levels);
    const int full = qtx::mip_level_count(w, h, 0);
    if (levels > full) return fail(QTX_EINVAL, "a %d x %d texture has %d mip levels, not %d", w, h, full, levels);
    ws_bytes = qtx::mip_plan(w, h, levels, p);
    return QTX_OK;
}

size_t align256(size_t n) { return (n + 255) & ~(size_t)255; }


}  // namespace

extern "C" {

int32_t qtx_mip_count(int32_t width, int32_t height, int32_t requested) {
    if (width <= 0 || height <= 0) return 0;
    return qtx::mip_level_count(width, height, requested);
}

size_t qtx_mips_workspace_bytes(int32_t width, int32_t height, int32_t levels) {
    qtx::MipPlan p;
    size_t ws = 0;
    if (chain_plan(width, height, levels, p, ws) != QTX_OK) return 0;
    return ws;
}

qtx_status qtx_mips_device(qtx_ctx *ctx, const uint8_t *d_rgba, int32_t width, int32_t height, int32_t levels,
                           uint8_t *const *d_levels, void *d_workspace, size_t workspace_bytes, void *stream) {
    qtx_status st = check_ctx(ctx);
    if (st != QTX_OK) return st;
    qtx::MipPlan p;
    size_t need = 0;
    if ((st = chain_plan(width, height, levels, p, need)) != QTX_OK) return st;
    if (!d_rgba || (levels > 1 && (!d_levels || !d_workspace))) return fail(QTX_EINVAL, "null buffer");
    if (workspace_bytes < need) return fail(QTX_EINVAL, "workspace holds %zu bytes, %zu needed", workspace_bytes, need);
    for (int m = 1; m < levels; m++) {
        if (!d_levels[m - 1]) return fail(QTX_EINVAL, "null buffer for level %d", m);
        p.dst[m] = d_levels[m - 1];
    }
    QTX_HIP(qtx::launch_mips(p, d_rgba, static_cast<uint8_t *>(d_workspace), (hipStream_t)stream));
    return QTX_OK;
}

qtx_status qtx_mips_host(qtx_ctx *ctx, const uint8_t *rgba, int32_t width, int32_t height, int32_t levels,
                         uint8_t *const *out_levels, const size_t *out_bytes) {
    qtx_status st = check_ctx(ctx);
    if (st != QTX_OK) return st;
    qtx::MipPlan p;
    size_t ws = 0;
    if ((st = chain_plan(width, height, levels, p, ws)) != QTX_OK) return st;
    if (!rgba || (levels > 1 && (!out_levels || !out_bytes))) return fail(QTX_EINVAL, "null buffer");
    std::vector<size_t> off(levels);
    size_t total = align256(raw_bytes(width, height));
    for (int m = 1; m < levels; m++) {
        const size_t n = raw_bytes(p.lw[m], p.lh[m]);
        if (!out_levels[m - 1]) return fail(QTX_EINVAL, "null buffer for level %d", m);
        if (out_bytes[m - 1] < n)
            return fail(QTX_EINVAL, "level %d: output buffer holds %zu bytes, %zu needed", m, out_bytes[m - 1], n);
        off[m] = total;
        total += align256(n);
    }
    const size_t ws_off = total;
    total += ws;
    std::lock_guard<std::mutex> lk(ctx->mu);
    if ((st = ensure(ctx->d_in, ctx->in_cap, total)) != QTX_OK) return st;
    QTX_HIP(ctx->stager.upload(ctx->d_in, rgba, raw_bytes(width, height), ctx->stream));
    for (int m = 1; m < levels; m++) p.dst[m] = ctx->d_in + off[m];
    QTX_HIP(qtx::launch_mips(p, ctx->d_in, ctx->d_in + ws_off, ctx->stream));
    for (int m = 1; m < levels; m++)
        QTX_HIP(ctx->stager.download(out_levels[m - 1], ctx->d_in + off[m], raw_bytes(p.lw[m], p.lh[m]), ctx->stream));
    QTX_HIP(hipStreamSynchronize(ctx->stream));
    return QTX_OK;
}

}  // extern "C"

namespace {

// The DDS chain encode: the base, already in ctx->d_in when `uploaded`, else
// copied from rgba; every level resampled and encoded as one batch; the blocks
// copied to out_blocks.  Holds ctx->mu.
qtx_status encode_chain(qtx_ctx *ctx, const qtx_job &tmpl, const uint8_t *rgba, int32_t width, int32_t height,
                        int32_t levels, uint8_t *const *out_blocks, const size_t *out_bytes, bool uploaded) {
    qtx::MipPlan p;
    size_t ws = 0;
    qtx_status st = chain_plan(width, height, levels, p, ws);
    if (st != QTX_OK) return st;
    // layout: base | levels 1.. | workspace | encoded levels
    std::vector<size_t> img_off(levels), blk_off(levels);
    std::vector<qtx_job> jobs(levels, tmpl);
    std::vector<qtx::bt::Prepared> prep(levels);
    size_t total = 0;
    for (int m = 0; m < levels; m++) {
        img_off[m] = total;
        total += align256(raw_bytes(p.lw[m], p.lh[m]));
    }
    const size_t ws_off = total;
    total += align256(ws);
    for (int m = 0; m < levels; m++) {
        jobs[m].width = p.lw[m];
        jobs[m].height = p.lh[m];
        jobs[m].src = rgba;  // validated below; replaced by device pointers
        jobs[m].dst = out_blocks[m];
        if ((st = qtx::bt::prepare(jobs[m], (size_t)m, prep[m])) != QTX_OK) return st;
        if (!out_blocks[m]) return fail(QTX_EINVAL, "null buffer for level %d", m);
        if (out_bytes[m] < prep[m].dst_bytes)
            return fail(QTX_EINVAL, "level %d: output buffer holds %zu bytes, %zu needed", m, out_bytes[m],
                        prep[m].dst_bytes);
        blk_off[m] = total;
        total += align256(prep[m].dst_bytes);
    }
    if (uploaded) {
        if (ctx->in_cap < total) return fail(QTX_EINVAL, "staging holds %zu bytes, %zu needed", ctx->in_cap, total);
    } else {
        if ((st = ensure(ctx->d_in, ctx->in_cap, total)) != QTX_OK) return st;
        QTX_HIP(ctx->stager.upload(ctx->d_in, rgba, raw_bytes(width, height), ctx->stream));
    }
    uint8_t *d = ctx->d_in;
    for (int m = 1; m < levels; m++) p.dst[m] = d + img_off[m];
    QTX_HIP(qtx::launch_mips(p, d, d + ws_off, ctx->stream));
    for (int m = 0; m < levels; m++) {
        jobs[m].src = d + img_off[m];
        jobs[m].dst = d + blk_off[m];
    }
    std::vector<qtx::BatchJob> table;  // read by the copy to the device until the stream is synchronised
    st = qtx::bt::run_once(ctx, jobs, ctx->stream, table);
    if (st == QTX_OK) {  // the levels' blocks lie in one range: one staged transfer
        std::vector<qtx::rt::HostPart> parts(levels);
        for (int m = 0; m < levels; m++) parts[m] = {out_blocks[m], blk_off[m] - blk_off[0], prep[m].dst_bytes};
        const size_t n = blk_off[levels - 1] - blk_off[0] + prep[levels - 1].dst_bytes;
        const hipError_t e = ctx->stager.download_parts(parts.data(), parts.size(), d + blk_off[0], n, ctx->stream);
        if (e != hipSuccess) st = hip_fail(e, "download");
    }
    const hipError_t e = hipStreamSynchronize(ctx->stream);
    if (st == QTX_OK && e != hipSuccess) st = hip_fail(e, "hipStreamSynchronize");
    return st;
}

// Bytes the chain encode stages for a chain whose levels are BC3 (the larger format).
size_t chain_stage_bytes(int32_t width, int32_t height, int32_t levels) {
    qtx::MipPlan p;
    size_t ws = 0;
    if (chain_plan(width, height, levels, p, ws) != QTX_OK) return 0;
    size_t total = align256(ws);
    for (int m = 0; m < levels; m++)
        total += align256(raw_bytes(p.lw[m], p.lh[m])) + align256(block_bytes(p.lw[m], p.lh[m], 16));
    return total;
}

// *has_alpha for the texture in device memory at d_rgba (16-byte aligned), on
// ctx->stream (waits for it).  Holds ctx->mu.
qtx_status alpha_flag(qtx_ctx *ctx, const uint8_t *d_rgba, uint64_t texels, hipStream_t stream, int32_t *has_alpha) {
    if (!ctx->d_flag) QTX_HIP(hipMalloc(&ctx->d_flag, 256));
    QTX_HIP(hipMemsetAsync(ctx->d_flag, 0, sizeof(uint32_t), stream));
    QTX_HIP(qtx::launch_has_alpha(d_rgba, texels, ctx->d_flag, stream));
    uint32_t flag = 0;
    QTX_HIP(hipMemcpyAsync(&flag, ctx->d_flag, sizeof flag, hipMemcpyDeviceToHost, stream));
    QTX_HIP(hipStreamSynchronize(stream));
    *has_alpha = flag != 0;
    return QTX_OK;
}

}  // namespace

extern "C" {

qtx_status qtx_mips_encode_host(qtx_ctx *ctx, const qtx_job *tmpl, const uint8_t *rgba, int32_t width, int32_t height,
                                int32_t levels, uint8_t *const *out_blocks, const size_t *out_bytes) {
    qtx_status st = check_ctx(ctx);
    if (st != QTX_OK) return st;
    if (!tmpl) return fail(QTX_EINVAL, "null job template");
    if (tmpl->op != QTX_OP_ENCODE) return fail(QTX_EINVAL, "the job template must encode");
    if (!rgba || !out_blocks || !out_bytes) return fail(QTX_EINVAL, "null buffer");
    std::lock_guard<std::mutex> lk(ctx->mu);
    return encode_chain(ctx, *tmpl, rgba, width, height, levels, out_blocks, out_bytes, false);
}

qtx_status qtx_mips_encode_auto_host(qtx_ctx *ctx, const qtx_bc1_params *bc1, const uint8_t *rgba, int32_t width,
                                     int32_t height, int32_t levels, uint8_t *const *out_blocks, const size_t *out_bytes,
                                     int32_t *format) {
    qtx_status st = check_ctx(ctx);
    if (st != QTX_OK) return st;
    if (!bc1 || !rgba || !out_blocks || !out_bytes || !format) return fail(QTX_EINVAL, "null pointer");
    qtx::MipPlan p;
    size_t ws = 0;
    if ((st = chain_plan(width, height, levels, p, ws)) != QTX_OK) return st;
    std::lock_guard<std::mutex> lk(ctx->mu);
    if ((st = ensure(ctx->d_in, ctx->in_cap, chain_stage_bytes(width, height, levels))) != QTX_OK) return st;
    QTX_HIP(ctx->stager.upload(ctx->d_in, rgba, raw_bytes(width, height), ctx->stream));
    int32_t has_alpha = 0;
    if ((st = alpha_flag(ctx, ctx->d_in, (uint64_t)width * (uint64_t)height, ctx->stream, &has_alpha)) != QTX_OK)
        return st;
    qtx_job j{};
    j.op = QTX_OP_ENCODE;
    j.format = has_alpha ? QTX_FMT_BC3 : QTX_FMT_BC1;
    j.bc1 = *bc1;
    if (has_alpha) j.bc1.color_mode = QTX_COLOR_FOUR;  // BC3Encoder(level): a FourColor BC1 encoder
    *format = j.format;
    return encode_chain(ctx, j, rgba, width, height, levels, out_blocks, out_bytes, true);
}

qtx_status qtx_has_alpha_device(qtx_ctx *ctx, const uint8_t *d_rgba, int32_t width, int32_t height, void *stream,
                                int32_t *has_alpha) {
    qtx_status st = check_ctx(ctx);
    if (st == QTX_OK) st = check_dims(width, height);
    if (st != QTX_OK) return st;
    if (!d_rgba || !has_alpha) return fail(QTX_EINVAL, "null pointer");
    if (reinterpret_cast<uintptr_t>(d_rgba) % 16) return fail(QTX_EINVAL, "texture not 16-byte aligned");
    std::lock_guard<std::mutex> lk(ctx->mu);
    return alpha_flag(ctx, d_rgba, (uint64_t)width * (uint64_t)height, (hipStream_t)stream, has_alpha);
}

qtx_status qtx_has_alpha_host(qtx_ctx *ctx, const uint8_t *rgba, int32_t width, int32_t height, int32_t *has_alpha) {
    qtx_status st = check_ctx(ctx);
    if (st == QTX_OK) st = check_dims(width, height);
    if (st != QTX_OK) return st;
    if (!rgba || !has_alpha) return fail(QTX_EINVAL, "null pointer");
    std::lock_guard<std::mutex> lk(ctx->mu);
    if ((st = ensure(ctx->d_in, ctx->in_cap, raw_bytes(width, height))) != QTX_OK) return st;
    QTX_HIP(ctx->stager.upload(ctx->d_in, rgba, raw_bytes(width, height), ctx->stream));
    return alpha_flag(ctx, ctx->d_in, (uint64_t)width * (uint64_t)height, ctx->stream, has_alpha);
}

qtx_status qtx_swizzle_bc3nm_device(qtx_ctx *ctx, uint8_t *d_rgba, int32_t width, int32_t height, void *stream) {
    qtx_status st = check_ctx(ctx);
    if (st == QTX_OK) st = check_dims(width, height);
    if (st != QTX_OK) return st;
    if (!d_rgba) return fail(QTX_EINVAL, "null buffer");
    QTX_HIP(qtx::launch_swizzle_bc3nm(d_rgba, (uint64_t)width * (uint64_t)height, (hipStream_t)stream));
    return QTX_OK;
}

qtx_status qtx_swizzle_bc3nm_host(qtx_ctx *ctx, const uint8_t *rgba, int32_t width, int32_t height, uint8_t *out) {
    qtx_status st = check_ctx(ctx);
    if (st == QTX_OK) st = check_dims(width, height);
    if (st != QTX_OK) return st;
    if (!rgba || !out) return fail(QTX_EINVAL, "null buffer");
    std::lock_guard<std::mutex> lk(ctx->mu);
    const size_t n = raw_bytes(width, height);
    if ((st = ensure(ctx->d_in, ctx->in_cap, n)) != QTX_OK) return st;
    QTX_HIP(ctx->stager.upload(ctx->d_in, rgba, n, ctx->stream));
    QTX_HIP(qtx::launch_swizzle_bc3nm(ctx->d_in, (uint64_t)width * (uint64_t)height, ctx->stream));
    QTX_HIP(ctx->stager.download(out, ctx->d_in, n, ctx->stream));
    QTX_HIP(hipStreamSynchronize(ctx->stream));
    return QTX_OK;
}

}  // extern "C"
