// kernels.hip -- gfx950 kernels of the S3TC path and their launchers.
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

#include "bc1_device.h"
#include "kernels.h"

namespace qtx {
namespace dev {

// ---------------------------------------------------------------------------
// Gather of one 4x4 block by one lane per texel (lane i of the group loads
// texel i) then shared across the group.
__device__ __forceinline__ uint32_t load_texel(const uint8_t *__restrict__ rgba, int w, int h, int bx, int by, int t) {
    int x = bx * 4 + (t & 3), y = by * 4 + (t >> 2);
    if (x >= w) x %= w;
    if (y >= h) y %= h;
    return *reinterpret_cast<const uint32_t *>(rgba + 4 * ((size_t)y * (size_t)w + (size_t)x));
}

// BC4Encoder::EncodeBlock (quicktex/s3tc/bc4/BC4Encoder.cpp:30-71); v[i] in [0,255].
__device__ __forceinline__ uint2 bc4_encode_values(const uint32_t v[16]) {
    uint32_t mn = 255, mx = 0;
#pragma unroll
    for (int i = 0; i < 16; i++) {
        mn = min(mn, v[i]);
        mx = max(mx, v[i]);
    }
    if (mx == mn) return make_uint2(mn | (mn << 8), 0u);
    const int bias = 4 - (int)mn * 14;
    const int delta = (int)(mx - mn);
    uint64_t bits = 0;
#pragma unroll
    for (int i = 0; i < 16; i++) {
        const int value = (int)v[i] * 14;
        int level = 0;
#pragma unroll
        for (int c = 0; c < 7; c++) level += value >= delta * (1 + 2 * c) - bias;
        const uint64_t s = (0x02345671u >> (4 * level)) & 7u;  // {1,7,6,5,4,3,2,0}
        bits |= s << (3 * i);
    }
    return make_uint2(mx | (mn << 8) | ((uint32_t)(bits & 0xFFFFu) << 16), (uint32_t)(bits >> 16));
}

// BC4Block values (quicktex/s3tc/bc4/BC4Block.cpp:44-64) and selector decode.
__device__ __forceinline__ uint32_t bc4_value(uint32_t a0, uint32_t a1, uint32_t s) {
    if (a0 <= a1) {  // 6-value
        if (s == 0) return a0;
        if (s == 1) return a1;
        if (s == 6) return 0;
        if (s == 7) return 255;
        return (a0 * (6 - s) + a1 * (s - 1)) / 5;
    }
    if (s == 0) return a0;
    if (s == 1) return a1;
    return (a0 * (8 - s) + a1 * (s - 1)) / 7;
}

// ---------------------------------------------------------------------------
template <int INTERP, bool BC3>
__global__ __launch_bounds__(256, 4) void k_bc1_encode(const uint8_t *__restrict__ rgba, int w, int h, int bw, int nblocks,
                                                       uint8_t *__restrict__ out, const DeviceTables *__restrict__ T,
                                                       BC1Settings s) {
    load_lds_tables<INTERP>(T, s);
    __syncthreads();
    // workgroups stride over chunks of 16 blocks; each group of 16 lanes owns one
    // block of the chunk and only touches its own 16 LDS texel words
    const int nchunks = (nblocks + 15) / 16;
    const int group = (int)(threadIdx.x / G);
    for (int chunk = (int)blockIdx.x; chunk < nchunks; chunk += (int)gridDim.x) {
        const int gid = chunk * 16 + group;
        if (gid >= nblocks) break;
        const int bx = gid % bw, by = gid / bw;
        const uint32_t px = load_texel(rgba, w, h, bx, by, glane());
        L().tex[threadIdx.x] = px;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        const uint32_t rgb0 = px & 0x00FFFFFFu;
        const int pp = (int)gsum(dot4(rgb0, rgb0));
        const uint32_t ppa = gsum(dot4(px, px));
        const Ctx cx{T, px, pp, ppa};
        const uint2 colour = encode_bc1<INTERP>(cx);
        uint2 alpha = make_uint2(0, 0);
        if (BC3) {
            uint32_t a[16];
#pragma unroll
            for (int i = 0; i < 16; i++) a[i] = tex(i) >> 24;
            alpha = bc4_encode_values(a);
        }
        if (glane() == 0) {
            if (BC3) reinterpret_cast<uint4 *>(out)[gid] = make_uint4(alpha.x, alpha.y, colour.x, colour.y);
            else reinterpret_cast<uint2 *>(out)[gid] = colour;
        }
        // the next chunk's texels overwrite this group's LDS words: keep the reads above them
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
    }
}

// BC4 (one channel) or BC5 (two channels): one lane per block.
template <bool BC5>
__global__ __launch_bounds__(256) void k_bc4_encode(const uint8_t *__restrict__ rgba, int w, int h, int bw, int nblocks,
                                                    int ch0, int ch1, uint8_t *__restrict__ out) {
    const int gid = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (gid >= nblocks) return;
    const int bx = gid % bw, by = gid / bw;
    uint32_t px[16];
    if ((w & 3) == 0 && bx * 4 + 3 < w && by * 4 + 3 < h) {
#pragma unroll
        for (int y = 0; y < 4; y++) {
            const uint4 row = *reinterpret_cast<const uint4 *>(rgba + 4 * ((size_t)(by * 4 + y) * w + bx * 4));
            px[4 * y + 0] = row.x; px[4 * y + 1] = row.y; px[4 * y + 2] = row.z; px[4 * y + 3] = row.w;
        }
    } else {
#pragma unroll
        for (int t = 0; t < 16; t++) px[t] = load_texel(rgba, w, h, bx, by, t);
    }
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

template <int INTERP, int KIND>
__global__ __launch_bounds__(256) void k_decode(const uint8_t *__restrict__ blocks, int w, int h, int bw, int bh,
                                                int p0, int p1, uint8_t *__restrict__ rgba) {
    const int gid = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (gid >= bw * bh) return;
    const int bx = gid % bw, by = gid / bw;
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
        uint32_t bc[4];
        bc[0] = pal[0] | 0xFF000000u;
        bc[1] = (three ? pal[2] : pal[3]) | 0xFF000000u;
        bc[2] = pal[1] | 0xFF000000u;
        bc[3] = three ? 0u : (pal[2] | 0xFF000000u);
        const bool write_alpha = (KIND == DEC_BC1) && p0 != 0;
#pragma unroll
        for (int i = 0; i < 16; i++) {
            const uint32_t s = (cb.y >> (2 * i)) & 3u;
            const uint32_t c = s == 0 ? bc[0] : (s == 1 ? bc[1] : (s == 2 ? bc[2] : bc[3]));
            px[i] = write_alpha ? c : ((c & 0x00FFFFFFu) | 0xFF000000u);
        }
        if (KIND == DEC_BC3) {
            const uint32_t a0 = ab.x & 0xFFu, a1 = (ab.x >> 8) & 0xFFu;
            const uint64_t bits = (uint64_t)(ab.x >> 16) | ((uint64_t)ab.y << 16);
#pragma unroll
            for (int i = 0; i < 16; i++) {
                const uint32_t v = bc4_value(a0, a1, (uint32_t)(bits >> (3 * i)) & 7u);
                px[i] = (px[i] & 0x00FFFFFFu) | (v << 24);
            }
        }
    } else {
        const int nch = KIND == DEC_BC5 ? 2 : 1;
#pragma unroll
        for (int k = 0; k < nch; k++) {
            const uint2 q = reinterpret_cast<const uint2 *>(blocks)[gid * nch + k];
            const int ch = k == 0 ? p0 : p1;
            const uint32_t a0 = q.x & 0xFFu, a1 = (q.x >> 8) & 0xFFu;
            const uint64_t bits = (uint64_t)(q.x >> 16) | ((uint64_t)q.y << 16);
            const uint32_t sh = 8u * (uint32_t)ch;
#pragma unroll
            for (int i = 0; i < 16; i++) {
                const uint32_t v = bc4_value(a0, a1, (uint32_t)(bits >> (3 * i)) & 7u);
                px[i] = (px[i] & ~(0xFFu << sh)) | (v << sh);
            }
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
                *reinterpret_cast<uint4 *>(row) = make_uint4(px[4 * y], px[4 * y + 1], px[4 * y + 2], px[4 * y + 3]);
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

}  // namespace dev

// ---------------------------------------------------------------------------
// Launchers
namespace {
constexpr int kThreads = 256;
constexpr size_t kBc1Lds = sizeof(dev::Lds);
// persistent workgroups for the BC1/BC3 encoder: 256 CUs x 4 resident workgroups
constexpr int kBc1MaxGrid = 256 * 4;
inline unsigned grid_for(long long items, int per_block) { return (unsigned)((items + per_block - 1) / per_block); }
}  // namespace

hipError_t launch_bc1_encode(const uint8_t *d_rgba, int w, int h, uint8_t *d_out, const DeviceTables *d_tables,
                             const BC1Settings &s, int interp, bool bc3, hipStream_t stream) {
    const int bw = (w + 3) / 4, bh = (h + 3) / 4;
    const int nblocks = bw * bh;
    const dim3 grid((unsigned)std::min<long long>(((long long)nblocks + 15) / 16, kBc1MaxGrid)), block(kThreads);
#define QTX_BC1_CASE(I)                                                                                             \
    case I:                                                                                                         \
        if (bc3)                                                                                                    \
            hipLaunchKernelGGL((dev::k_bc1_encode<I, true>), grid, block, kBc1Lds, stream, d_rgba, w, h, bw, nblocks, d_out, \
                               d_tables, s);                                                                        \
        else                                                                                                        \
            hipLaunchKernelGGL((dev::k_bc1_encode<I, false>), grid, block, kBc1Lds, stream, d_rgba, w, h, bw, nblocks,      \
                               d_out, d_tables, s);                                                                 \
        break;
    switch (interp) {
        QTX_BC1_CASE(QTX_INTERP_IDEAL)
        QTX_BC1_CASE(QTX_INTERP_IDEAL_ROUND)
        QTX_BC1_CASE(QTX_INTERP_NVIDIA)
        QTX_BC1_CASE(QTX_INTERP_AMD)
        default: return hipErrorInvalidValue;
    }
#undef QTX_BC1_CASE
    return hipGetLastError();
}

hipError_t launch_bc4_encode(const uint8_t *d_rgba, int w, int h, uint8_t *d_out, int ch0, int ch1, bool bc5,
                             hipStream_t stream) {
    const int bw = (w + 3) / 4, bh = (h + 3) / 4;
    const int nblocks = bw * bh;
    const dim3 grid(grid_for(nblocks, kThreads)), block(kThreads);
    if (bc5)
        hipLaunchKernelGGL((dev::k_bc4_encode<true>), grid, block, 0, stream, d_rgba, w, h, bw, nblocks, ch0, ch1, d_out);
    else
        hipLaunchKernelGGL((dev::k_bc4_encode<false>), grid, block, 0, stream, d_rgba, w, h, bw, nblocks, ch0, ch1, d_out);
    return hipGetLastError();
}

hipError_t launch_decode(int format, const uint8_t *d_blocks, int w, int h, uint8_t *d_rgba, int interp, int p0, int p1,
                         hipStream_t stream) {
    const int bw = (w + 3) / 4, bh = (h + 3) / 4;
    const dim3 grid(grid_for((long long)bw * bh, kThreads)), block(kThreads);
#define QTX_DEC(I, K) \
    hipLaunchKernelGGL((dev::k_decode<I, K>), grid, block, 0, stream, d_blocks, w, h, bw, bh, p0, p1, d_rgba)
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

}  // namespace qtx
