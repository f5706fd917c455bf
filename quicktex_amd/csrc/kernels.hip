// kernels.hip -- gfx950 kernels of the S3TC path other than the BC1/BC3 encoders
// (BC4/BC5 encode, the decoders, the batch's light kernel), and every launcher:
// the BC1/BC3 encode kernels are instantiated per interpolator (bc1_kernels.h).
//
// Layouts (quicktex/Texture.h:73-87, 158-162): raw textures are row-major RGBA8
// with no row padding; block textures are row-major blocks of 8 (BC1/BC4) or
// 16 (BC3/BC5) bytes.
//
// Edge rules (quicktex/Texture.h:89-134):
//   encode gathers texel (tx, ty) of block (bx, by) from ((4bx+tx) mod w, (4by+ty) mod h);
//   decode writes blocks in row-major order, texels x-outer / y-inner, so a
//   texture pixel keeps the value of the last block texel that maps onto it.
//   Per axis that is the last texel of the last block if the last block wraps
//   onto the pixel, else the pixel's own position; decode writes only those
//   texels, so no two lanes write one pixel.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "bc1_lane.h"
#include "bc1_launch.h"
#include "kernels.h"
#include "kernels_dev.h"

namespace qtx {
namespace dev {

// The eight values of a BC4 block (quicktex/s3tc/bc4/BC4Block.cpp:44-64) as
// bytes, value s in byte s of (lo, hi): both rules computed, one kept, so lanes
// of either kind do not diverge.
__device__ __forceinline__ uint2 bc4_palette(uint32_t a0, uint32_t a1) {
    uint32_t six[4], eight[6];
#pragma unroll
    for (uint32_t s = 2; s < 6; s++) six[s - 2] = (a0 * (6 - s) + a1 * (s - 1)) / 5;
#pragma unroll
    for (uint32_t s = 2; s < 8; s++) eight[s - 2] = (a0 * (8 - s) + a1 * (s - 1)) / 7;
    const uint32_t base = a0 | (a1 << 8);
    return a0 <= a1 ? make_uint2(base | (six[0] << 16) | (six[1] << 24), six[2] | (six[3] << 8) | 0xFF000000u)
                    : make_uint2(base | (eight[0] << 16) | (eight[1] << 24),
                                 eight[2] | (eight[3] << 8) | (eight[4] << 16) | (eight[5] << 24));
}

// The texels' channel ch (byte sh / 8) set from a BC4 block (a0, a1, 48
// selector bits): each value picked out of the palette bytes by v_perm_b32 and
// placed at its byte in the same step (selector byte 0x0C gives zero).
__device__ __forceinline__ void bc4_decode_into(uint32_t px[16], uint32_t a0, uint32_t a1, uint64_t bits, uint32_t sh) {
    const uint2 pal = bc4_palette(a0, a1);
    const uint32_t keep = ~(0xFFu << sh), zero = 0x0C0C0C0Cu & keep;
#pragma unroll
    for (int i = 0; i < 16; i++) {
        const uint32_t sel = zero | (((uint32_t)(bits >> (3 * i)) & 7u) << sh);
        px[i] = (px[i] & keep) | __builtin_amdgcn_perm(pal.y, pal.x, sel);
    }
}

// The 16 texels of block (bx, by) (edge blocks wrap like the reference's texel reads).
__device__ __forceinline__ void load_block_texels(const uint8_t *__restrict__ rgba, int w, int h, int bx, int by,
                                                  uint32_t px[16]) {
    if ((w & 3) == 0 && bx * 4 + 3 < w && by * 4 + 3 < h) {
        // every texel is read once: non-temporal loads keep it out of the caches
        typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
#pragma unroll
        for (int y = 0; y < 4; y++) {
            const u32x4 row = __builtin_nontemporal_load(
                reinterpret_cast<const u32x4 *>(rgba + 4 * ((size_t)(by * 4 + y) * w + bx * 4)));
            px[4 * y + 0] = row.x; px[4 * y + 1] = row.y; px[4 * y + 2] = row.z; px[4 * y + 3] = row.w;
        }
    } else {
#pragma unroll
        for (int t = 0; t < 16; t++) px[t] = load_texel(rgba, w, h, bx, by, t);
    }
}

// BC4 (one channel) or BC5 (two channels) of block gid: one lane per block.
template <bool BC5>
__device__ __forceinline__ void encode_bc45_block(const uint8_t *__restrict__ rgba, int w, int h, int gid, int bx, int by,
                                                  int ch0, int ch1, uint8_t *__restrict__ out) {
    uint32_t px[16];
    load_block_texels(rgba, w, h, bx, by, px);
    uint32_t v[16];
#pragma unroll
    for (int i = 0; i < 16; i++) v[i] = chn(px[i], ch0);
    const uint2 b0 = bc4_encode_values(v);
    if (BC5) {
#pragma unroll
        for (int i = 0; i < 16; i++) v[i] = chn(px[i], ch1);
        const uint2 b1 = bc4_encode_values(v);
        reinterpret_cast<uint4 *>(out)[gid] = make_uint4(b0.x, b0.y, b1.x, b1.y);
    } else {
        reinterpret_cast<uint2 *>(out)[gid] = b0;
    }
}

template <bool BC5>
__global__ __launch_bounds__(256) void k_bc4_encode(const uint8_t *__restrict__ rgba, int w, int h, int bw, RowDiv rd,
                                                    int nblocks, int ch0, int ch1, uint8_t *__restrict__ out) {
    const int gid = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (gid >= nblocks) return;
    int bx, by;
    block_xy((uint32_t)gid, (uint32_t)bw, rd, bx, by);
    encode_bc45_block<BC5>(rgba, w, h, gid, bx, by, ch0, ch1, out);
}

// ---------------------------------------------------------------------------
// Decoders: one lane per block.
enum DecKind { DEC_BC1 = 1, DEC_BC3 = 3, DEC_BC4 = 4, DEC_BC5 = 5 };

// true if texel t of block b is the final writer of its pixel along an axis of
// n pixels and nb blocks.
__device__ __forceinline__ bool axis_winner(int b, int t, int n, int nb) {
    const int X = (b * 4 + t) % n;
    const int last = nb - 1;
#pragma unroll
    for (int tt = 3; tt >= 0; tt--) {
        if ((last * 4 + tt) % n == X) return b == last && t == tt;
    }
    return b == (X >> 2) && t == (X & 3);
}

// A 16-byte row of decoded texels, stored non-temporally: the decoders write 8x
// the bytes they read and nothing reads the texels back, so the stores stream
// past the caches (measured 2.1x faster on a 4096x4096 decode than plain stores).
__device__ __forceinline__ void store_row_nt(uint8_t *row, uint32_t a, uint32_t b, uint32_t c, uint32_t d) {
    typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
    const u32x4 v = {a, b, c, d};
    __builtin_nontemporal_store(v, reinterpret_cast<u32x4 *>(row));
}

// Decode of block gid of a bw x bh block texture into the w x h RGBA8 texture.
template <int INTERP, int KIND>
__device__ __forceinline__ void decode_block(const uint8_t *__restrict__ blocks, int w, int h, int bw, int bh, int p0,
                                             int p1, uint8_t *__restrict__ rgba, int gid, int bx, int by) {
    uint32_t px[16];
#pragma unroll
    for (int i = 0; i < 16; i++) px[i] = 0xFF000000u;  // Color() = (0, 0, 0, 255)
    if (KIND == DEC_BC1 || KIND == DEC_BC3) {
        uint2 cb;
        uint2 ab = make_uint2(0, 0);
        if (KIND == DEC_BC3) {
            const uint4 q = reinterpret_cast<const uint4 *>(blocks)[gid];
            ab = make_uint2(q.x, q.y);
            cb = make_uint2(q.z, q.w);
        } else {
            cb = reinterpret_cast<const uint2 *>(blocks)[gid];
        }
        // BC1Decoder::DecodeBlock (BC1Decoder.cpp:31-52) via Interpolate565BC1 (Interpolator.cpp:52-55)
        const uint32_t c0 = cb.x & 0xFFFFu, c1 = cb.x >> 16;
        const bool three = (KIND == DEC_BC1) && (c1 >= c0);
        const uint32_t lo = rgb(c0 >> 11, (c0 >> 5) & 63u, c0 & 31u);
        const uint32_t hi = rgb(c1 >> 11, (c1 >> 5) & 63u, c1 & 31u);
        uint32_t pal[4];
        palette<INTERP>(lo, hi, three, pal);
        // back to BC1 selector order {c0, c1, c2, c3}; alpha 255 except 3-colour black
        // (3-colour black keeps alpha 0 only when alpha is written; BC3 takes alpha
        // from its alpha block below)
        const bool write_alpha = (KIND == DEC_BC1) && p0 != 0;
        const uint32_t bc0 = pal[0] | 0xFF000000u;
        const uint32_t bc1 = (three ? pal[2] : pal[3]) | 0xFF000000u;
        const uint32_t bc2 = pal[1] | 0xFF000000u;
        const uint32_t bc3 = three ? (write_alpha ? 0u : 0xFF000000u) : (pal[2] | 0xFF000000u);
        {
            // byte planes: R holds the red of colours 0-3 in bytes 0-3, and so on; a
            // texel is then two v_perm_b32 (red and green, blue and alpha) of its selector
            // (BC3: alpha from the alpha block below)
            const uint32_t lo01 = __builtin_amdgcn_perm(bc1, bc0, 0x05010400u);
            const uint32_t hi01 = __builtin_amdgcn_perm(bc1, bc0, 0x07030602u);
            const uint32_t lo23 = __builtin_amdgcn_perm(bc3, bc2, 0x05010400u);
            const uint32_t hi23 = __builtin_amdgcn_perm(bc3, bc2, 0x07030602u);
            const uint32_t R = __builtin_amdgcn_perm(lo23, lo01, 0x05040100u);
            const uint32_t G = __builtin_amdgcn_perm(lo23, lo01, 0x07060302u);
            const uint32_t B = __builtin_amdgcn_perm(hi23, hi01, 0x05040100u);
            const uint32_t A = __builtin_amdgcn_perm(hi23, hi01, 0x07060302u);
#pragma unroll
            for (int i = 0; i < 16; i++) {
                const uint32_t s = (cb.y >> (2 * i)) & 3u;
                const uint32_t sel = s * 0x101u + 0x0C0C0400u;  // bytes: s, s + 4, zero, zero
                px[i] = __builtin_amdgcn_perm(G, R, sel) | __builtin_amdgcn_perm(A, B, (sel << 16) | 0x0C0Cu);
            }
        }
        if (KIND == DEC_BC3)
            bc4_decode_into(px, ab.x & 0xFFu, (ab.x >> 8) & 0xFFu, (uint64_t)(ab.x >> 16) | ((uint64_t)ab.y << 16), 24u);
    } else {
        const int nch = KIND == DEC_BC5 ? 2 : 1;
        uint2 qs[2];
        if (KIND == DEC_BC5) {  // both channels' blocks in one 16-byte load
            const uint4 v = reinterpret_cast<const uint4 *>(blocks)[gid];
            qs[0] = make_uint2(v.x, v.y);
            qs[1] = make_uint2(v.z, v.w);
        } else {
            qs[0] = reinterpret_cast<const uint2 *>(blocks)[gid];
        }
#pragma unroll
        for (int k = 0; k < nch; k++) {
            const uint2 q = qs[k];
            const int ch = k == 0 ? p0 : p1;
            bc4_decode_into(px, q.x & 0xFFu, (q.x >> 8) & 0xFFu, (uint64_t)(q.x >> 16) | ((uint64_t)q.y << 16),
                            8u * (uint32_t)ch);
        }
    }
    // scatter
    const int wrapx = bw * 4 - w, wrapy = bh * 4 - h;
    const bool fast_x = wrapx == 0 || (bx != 0 && bx != bw - 1);
    const bool fast_y = wrapy == 0 || (by != 0 && by != bh - 1);
    if (fast_x && fast_y) {
#pragma unroll
        for (int y = 0; y < 4; y++) {
            uint8_t *row = rgba + 4 * ((size_t)(by * 4 + y) * w + bx * 4);
            if ((w & 3) == 0) {
                store_row_nt(row, px[4 * y], px[4 * y + 1], px[4 * y + 2], px[4 * y + 3]);
            } else {
#pragma unroll
                for (int x = 0; x < 4; x++) reinterpret_cast<uint32_t *>(row)[x] = px[4 * y + x];
            }
        }
    } else {
        bool wx[4], wy[4];
#pragma unroll
        for (int t = 0; t < 4; t++) {
            wx[t] = axis_winner(bx, t, w, bw);
            wy[t] = axis_winner(by, t, h, bh);
        }
#pragma unroll
        for (int y = 0; y < 4; y++)
#pragma unroll
            for (int x = 0; x < 4; x++) {
                if (!(wx[x] && wy[y])) continue;
                const int X = (bx * 4 + x) % w, Y = (by * 4 + y) % h;
                *reinterpret_cast<uint32_t *>(rgba + 4 * ((size_t)Y * w + X)) = px[4 * y + x];
            }
    }
}

template <int INTERP, int KIND>
__global__ __launch_bounds__(256) void k_decode(const uint8_t *__restrict__ blocks, int w, int h, int bw, int bh,
                                                RowDiv rd, int p0, int p1, uint8_t *__restrict__ rgba) {
    const int gid = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (gid >= bw * bh) return;
    int bx, by;
    block_xy((uint32_t)gid, (uint32_t)bw, rd, bx, by);
    decode_block<INTERP, KIND>(blocks, w, h, bw, bh, p0, p1, rgba, gid, bx, by);
}

// ---------------------------------------------------------------------------
// Batch kernels.  A batch is a list of jobs (textures); each kernel of a batch
// owns a job array sorted by descending cost and a queue of work units numbered
// job by job.

// BC4/BC5 encode and every decoder of a batch: one lane per block, one unit of
// 256 blocks per workgroup.
template <int KIND>
__device__ __forceinline__ void decode_interp(const BatchJob &J, int gid, int bx, int by) {
    switch (J.interp) {
        case QTX_INTERP_IDEAL_ROUND:
            decode_block<QTX_INTERP_IDEAL_ROUND, KIND>(J.src, J.w, J.h, J.bw, J.bh, J.p0, J.p1, J.dst, gid, bx, by);
            break;
        case QTX_INTERP_NVIDIA:
            decode_block<QTX_INTERP_NVIDIA, KIND>(J.src, J.w, J.h, J.bw, J.bh, J.p0, J.p1, J.dst, gid, bx, by);
            break;
        case QTX_INTERP_AMD: decode_block<QTX_INTERP_AMD, KIND>(J.src, J.w, J.h, J.bw, J.bh, J.p0, J.p1, J.dst, gid, bx, by); break;
        default: decode_block<QTX_INTERP_IDEAL, KIND>(J.src, J.w, J.h, J.bw, J.bh, J.p0, J.p1, J.dst, gid, bx, by); break;
    }
}

// Workgroups stride over the units, so a launch may use fewer workgroups than
// units (a decode that writes host memory is bound by PCIe, not by the CUs).
__global__ __launch_bounds__(256) void k_light_batch(const BatchJob *__restrict__ jobs, int njobs, uint32_t nunits) {
    for (uint32_t u = blockIdx.x; u < nunits; u += gridDim.x) {
        const BatchJob &J = unit_job(jobs, njobs, u);
        const uint32_t gid = (u - J.unit_begin) * 256u + threadIdx.x;
        if (gid >= J.nblocks) continue;
        const int bx = (int)(gid % (uint32_t)J.bw), by = (int)(gid / (uint32_t)J.bw);
        switch (J.kind) {
            case BJ_ENC_BC4: encode_bc45_block<false>(J.src, J.w, J.h, (int)gid, bx, by, J.p0, J.p1, J.dst); break;
            case BJ_ENC_BC5: encode_bc45_block<true>(J.src, J.w, J.h, (int)gid, bx, by, J.p0, J.p1, J.dst); break;
            case BJ_DEC_BC1: decode_interp<DEC_BC1>(J, (int)gid, bx, by); break;
            case BJ_DEC_BC3: decode_interp<DEC_BC3>(J, (int)gid, bx, by); break;
            case BJ_DEC_BC4: decode_block<0, DEC_BC4>(J.src, J.w, J.h, J.bw, J.bh, J.p0, J.p1, J.dst, (int)gid, bx, by); break;
            case BJ_DEC_BC5: decode_block<0, DEC_BC5>(J.src, J.w, J.h, J.bw, J.bh, J.p0, J.p1, J.dst, (int)gid, bx, by); break;
            default: break;
        }
    }
}

// Writes zero to n words (a kernel, not a memset node, so that a caller's graph
// capture of the launch replays it faithfully).
__global__ __launch_bounds__(256) void k_zero_words(uint32_t *__restrict__ words, uint32_t n) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) words[i] = 0u;
}

}  // namespace dev

// ---------------------------------------------------------------------------
// Launchers
namespace {
constexpr int kThreads = 256;
// persistent workgroups for the BC1/BC3 encoder: 256 CUs x 4 resident workgroups
constexpr int kBc1MaxGrid = 256 * 4;
inline unsigned grid_for(long long items, int per_block) { return (unsigned)((items + per_block - 1) / per_block); }
}  // namespace

// A fit of the oracle's op accounting on the block mixture (levels 0-19): about 2k
// ops of metrics, endpoints and selectors, 500 per cluster-fit ordering and pass,
// and 780 per endpoint-search round, of which about 37 run before the search stops
// for lack of improvement.
double bc1_block_ops(const BC1Settings &s) {
    const double o4 = s.exhaustive ? kOrders4 : s.orderings4;
    const double o3 = s.exhaustive ? kOrders3 : s.orderings3;
    const bool three = (s.color_mode & QTX_COLOR_THREE) != 0;
    double ops = 2000.0 + s.cf_passes * (o4 * 500.0 + (three ? o3 * 500.0 + 1000.0 : 0.0));
    ops += 780.0 * (s.search_rounds < 37u ? (double)s.search_rounds : 37.0);
    return ops;
}

// Crossovers measured on MI355X (ThreeColorBlack, block mixture, Ideal): the static
// lane kernel (levels 0-9) wins from 32k blocks; the queued kernel, whose waves run
// one whole block per lane (a fixed cost of one block's latency), from about
// 65k + ops/2 blocks: 81k at level 10, 93k at level 13, 125k at level 16, 150k at
// level 18.  Level 19 (exhaustive cluster fit, about 1 M ops per block): the queued
// kernel takes 3.7 ms at any size up to 131k blocks (one block per lane) against
// 0.23-3.0 ms for the 16-lane encoder, 4.8 against 5.9 ms at 262k blocks and 12.0
// against 20.5 ms at 1 M: from kLaneFullBlocks (the formula gives 565k).
bool use_lane_encoder(const BC1Settings &s, uint64_t blocks, int path) {
    const int p = path != QTX_BC1_AUTO ? path : bc1_path_env();
    if (p == QTX_BC1_GROUP) return false;
    if (p == QTX_BC1_LANE) return true;
    if (blocks >= kLaneFullBlocks) return true;
    const double ops = bc1_block_ops(s);
    if (ops < kLaneCheapOps) return blocks >= kLaneMinBlocks;
    return (double)blocks >= kLaneQueueBaseBlocks + 0.5 * ops;
}

// The batch's lane kernel (k_lane_batch) deals no blocks and parks no searches:
// it wins from 32k blocks at settings up to 120 k ops/block, from 262k above.
bool use_lane_batch(const BC1Settings &s, uint64_t blocks, int path) {
    const int p = path != QTX_BC1_AUTO ? path : bc1_path_env();
    if (p == QTX_BC1_GROUP) return false;
    if (p == QTX_BC1_LANE) return dev::lane::lane_settings(s.exhaustive);
    if (!dev::lane::lane_settings(s.exhaustive)) return false;
    if (blocks >= kLaneFullBlocks) return true;
    return blocks >= kLaneMinBlocks && bc1_block_ops(s) <= kLaneBatchMaxOps;
}

int bc1_route(const BC1Settings &s, int w, int h, int path) {
    const uint64_t nblocks = (uint64_t)((w + 3) / 4) * (uint64_t)((h + 3) / 4);
    if (!use_lane_encoder(s, nblocks, path)) return kRouteGroup;
    return lane_queue_env() && bc1_block_ops(s) >= kLaneQueueMinOps ? kRouteQueue : kRouteLane;
}

bool wants_lane_queue(const BC1Settings &s, int w, int h, int path) {
    const int route = bc1_route(s, w, h, path);
    return route == kRouteQueue || (route == kRouteLane && lane_list_env() && bc1_block_ops(s) >= kLaneListMinOps);
}

size_t bc1_park_bytes(const BC1Settings &s, int w, int h, int path) {
    const size_t nblocks = (size_t)((w + 3) / 4) * (size_t)((h + 3) / 4);
    switch (bc1_route(s, w, h, path)) {
        case kRouteQueue: {
            const uint32_t stop = park_stop_env();
            return stop == 0 || s.search_rounds <= stop ? 0 : kParkRecordBytes * nblocks;
        }
        case kRouteLane:
            return lane_list_env() && bc1_block_ops(s) >= kLaneListMinOps
                       ? sizeof(uint32_t) * kLaneQueues * list_capacity((uint32_t)nblocks)
                       : 0;
        default: return 0;
    }
}

hipError_t launch_bc1_encode(const uint8_t *d_rgba, int w, int h, uint8_t *d_out, const DeviceTables *d_tables,
                             const BC1Settings &s, int interp, bool bc3, int path, uint32_t *d_queue, void *d_park,
                             int cu_count, hipStream_t stream) {
    const int bw = (w + 3) / 4, bh = (h + 3) / 4;
    const int nblocks = bw * bh;
    Bc1Launch a{};
    a.rgba = d_rgba;
    a.w = w;
    a.h = h;
    a.bw = bw;
    a.nblocks = nblocks;
    a.out = d_out;
    a.tables = d_tables;
    a.s = s;
    a.bc3 = bc3;
    a.stream = stream;
    const int route = bc1_route(s, w, h, path);
    if (route != kRouteGroup) {
        if (d_queue && route == kRouteQueue) {
            // persistent: one workgroup per CU (its LDS takes the CU's), of 12 waves, or
            // of 4 when the launch has at most a block per lane that way: a block then
            // runs on a wave alone on its SIMD (a third of the time it takes among 3)
            // and the blocks spread over every CU
            a.route = kRouteQueue;
            const int cus = std::max(cu_count, 1);
            a.threads = nblocks <= (long long)cus * kThreads && queue_small_env() ? kThreads : dev::lane::kQueueThreads;
            a.grid = (unsigned)std::min<long long>(grid_for(nblocks, a.threads), cus);
            // parking (and the drain launch after the queued kernel) only with a park buffer
            a.drain = d_park && bc1_park_bytes(s, w, h, path) > 0;
            a.stop = a.drain ? park_stop_env() : 0xFFFFFFFFu;
            a.drain_grid = (unsigned)std::max(cu_count, 1);
            a.fix = dev::lane::fixed_build(s);
            a.queue = d_queue;
            a.park = static_cast<uint4 *>(d_park);
        } else {
            a.route = kRouteLane;
            a.grid = grid_for(nblocks, kThreads);
            a.light = dev::lane::light_settings(s);
            // the block list (solid blocks stored first) when a queue set and list
            // words came with the launch
            if (d_queue && d_park && bc1_park_bytes(s, w, h, path) > 0) {
                a.queue = d_queue;
                a.list = static_cast<uint32_t *>(d_park);
            }
        }
    } else {
        a.route = kRouteGroup;
        a.grid = (unsigned)std::min<long long>(((long long)nblocks + 15) / 16, kBc1MaxGrid);
    }
    switch (interp) {
        case QTX_INTERP_IDEAL: return launch_bc1_interp<QTX_INTERP_IDEAL>(a);
        case QTX_INTERP_IDEAL_ROUND: return launch_bc1_interp<QTX_INTERP_IDEAL_ROUND>(a);
        case QTX_INTERP_NVIDIA: return launch_bc1_interp<QTX_INTERP_NVIDIA>(a);
        case QTX_INTERP_AMD: return launch_bc1_interp<QTX_INTERP_AMD>(a);
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_zero_words(uint32_t *d_words, uint32_t n, hipStream_t stream) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(dev::k_zero_words, dim3(std::min<unsigned>(grid_for(n, kThreads), 64u)), dim3(kThreads), 0,
                       stream, d_words, n);
    return hipGetLastError();
}

hipError_t launch_bc4_encode(const uint8_t *d_rgba, int w, int h, uint8_t *d_out, int ch0, int ch1, bool bc5,
                             hipStream_t stream) {
    const int bw = (w + 3) / 4, bh = (h + 3) / 4;
    const int nblocks = bw * bh;
    const dim3 grid(grid_for(nblocks, kThreads)), block(kThreads);
    const dev::RowDiv rd = dev::make_row_div((uint32_t)bw);
    if (bc5)
        hipLaunchKernelGGL((dev::k_bc4_encode<true>), grid, block, 0, stream, d_rgba, w, h, bw, rd, nblocks, ch0, ch1, d_out);
    else
        hipLaunchKernelGGL((dev::k_bc4_encode<false>), grid, block, 0, stream, d_rgba, w, h, bw, rd, nblocks, ch0, ch1, d_out);
    return hipGetLastError();
}

hipError_t launch_decode(int format, const uint8_t *d_blocks, int w, int h, uint8_t *d_rgba, int interp, int p0, int p1,
                         hipStream_t stream) {
    const int bw = (w + 3) / 4, bh = (h + 3) / 4;
    const dim3 grid(grid_for((long long)bw * bh, kThreads)), block(kThreads);
    const dev::RowDiv rd = dev::make_row_div((uint32_t)bw);
#define QTX_DEC(I, K) \
    hipLaunchKernelGGL((dev::k_decode<I, K>), grid, block, 0, stream, d_blocks, w, h, bw, bh, rd, p0, p1, d_rgba)
    switch (format) {
        case QTX_FMT_BC4: QTX_DEC(0, dev::DEC_BC4); break;
        case QTX_FMT_BC5: QTX_DEC(0, dev::DEC_BC5); break;
        case QTX_FMT_BC1:
        case QTX_FMT_BC3: {
            const bool b3 = format == QTX_FMT_BC3;
            switch (interp) {
                case QTX_INTERP_IDEAL: if (b3) QTX_DEC(QTX_INTERP_IDEAL, dev::DEC_BC3); else QTX_DEC(QTX_INTERP_IDEAL, dev::DEC_BC1); break;
                case QTX_INTERP_IDEAL_ROUND: if (b3) QTX_DEC(QTX_INTERP_IDEAL_ROUND, dev::DEC_BC3); else QTX_DEC(QTX_INTERP_IDEAL_ROUND, dev::DEC_BC1); break;
                case QTX_INTERP_NVIDIA: if (b3) QTX_DEC(QTX_INTERP_NVIDIA, dev::DEC_BC3); else QTX_DEC(QTX_INTERP_NVIDIA, dev::DEC_BC1); break;
                case QTX_INTERP_AMD: if (b3) QTX_DEC(QTX_INTERP_AMD, dev::DEC_BC3); else QTX_DEC(QTX_INTERP_AMD, dev::DEC_BC1); break;
                default: return hipErrorInvalidValue;
            }
            break;
        }
        default: return hipErrorInvalidValue;
    }
#undef QTX_DEC
    return hipGetLastError();
}

hipError_t launch_color_batch(int interp, const BatchJob *d_jobs, int njobs, uint32_t nunits, uint32_t *d_queue,
                              unsigned grid, const DeviceTables *d_tables, hipStream_t stream) {
    switch (interp) {
        case QTX_INTERP_IDEAL:
            return launch_bc1_batch_interp<QTX_INTERP_IDEAL>(false, d_jobs, njobs, nunits, d_queue, grid, d_tables, stream);
        case QTX_INTERP_IDEAL_ROUND:
            return launch_bc1_batch_interp<QTX_INTERP_IDEAL_ROUND>(false, d_jobs, njobs, nunits, d_queue, grid, d_tables,
                                                                   stream);
        case QTX_INTERP_NVIDIA:
            return launch_bc1_batch_interp<QTX_INTERP_NVIDIA>(false, d_jobs, njobs, nunits, d_queue, grid, d_tables, stream);
        case QTX_INTERP_AMD:
            return launch_bc1_batch_interp<QTX_INTERP_AMD>(false, d_jobs, njobs, nunits, d_queue, grid, d_tables, stream);
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_lane_batch(int interp, const BatchJob *d_jobs, int njobs, uint32_t nunits, unsigned grid,
                             const DeviceTables *d_tables, hipStream_t stream) {
    switch (interp) {
        case QTX_INTERP_IDEAL:
            return launch_bc1_batch_interp<QTX_INTERP_IDEAL>(true, d_jobs, njobs, nunits, nullptr, grid, d_tables, stream);
        case QTX_INTERP_IDEAL_ROUND:
            return launch_bc1_batch_interp<QTX_INTERP_IDEAL_ROUND>(true, d_jobs, njobs, nunits, nullptr, grid, d_tables,
                                                                   stream);
        case QTX_INTERP_NVIDIA:
            return launch_bc1_batch_interp<QTX_INTERP_NVIDIA>(true, d_jobs, njobs, nunits, nullptr, grid, d_tables, stream);
        case QTX_INTERP_AMD:
            return launch_bc1_batch_interp<QTX_INTERP_AMD>(true, d_jobs, njobs, nunits, nullptr, grid, d_tables, stream);
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_light_batch(const BatchJob *d_jobs, int njobs, uint32_t nunits, unsigned grid, hipStream_t stream) {
    hipLaunchKernelGGL(dev::k_light_batch, dim3(std::max(1u, std::min(grid, nunits))), dim3(kThreads), 0, stream, d_jobs,
                       njobs, nunits);
    return hipGetLastError();
}

}  // namespace qtx

