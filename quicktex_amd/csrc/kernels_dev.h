// kernels_dev.h -- device helpers shared by the kernel translation units:
// texel gathers with the reference's edge wrap, the BC4 block encoder, block
// stores and the batch's unit-to-job lookup.
//
// Layouts (quicktex/Texture.h:73-87, 158-162): raw textures are row-major RGBA8
// with no row padding; block textures are row-major blocks of 8 (BC1/BC4) or
// 16 (BC3/BC5) bytes.  Encode gathers texel (tx, ty) of block (bx, by) from
// ((4bx+tx) mod w, (4by+ty) mod h) (quicktex/Texture.h:89-134).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "bc1_device.h"
#include "qtx_internal.h"

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

// Division of a block index by the texture's width in blocks without a divide
// sequence (Granlund and Montgomery, "Division by invariant integers using
// multiplication", PLDI 1994, theorem 4.2 with N = 31 dividend bits): for
// 1 < bw < 2^31, l = ceil(log2 bw), m = floor(2^(31+l) / bw) + 1 (below 2^32) and
// s = l - 1, floor(gid / bw) = mulhi(gid, m) >> s for every gid < 2^31; m = 0
// stands for bw = 1.
struct RowDiv {
    uint32_t m, s;
};
inline RowDiv make_row_div(uint32_t bw) {
    if (bw <= 1) return RowDiv{0u, 0u};
    uint32_t l = 0;
    while ((1ull << l) < bw) l++;
    return RowDiv{(uint32_t)((1ull << (31 + l)) / bw + 1), l - 1};
}
// (bx, by) of block gid of a texture bw blocks wide
__device__ __forceinline__ void block_xy(uint32_t gid, uint32_t bw, RowDiv d, int &bx, int &by) {
    const uint32_t y = d.m ? __umulhi(gid, d.m) >> d.s : gid;
    by = (int)y;
    bx = (int)(gid - y * bw);
}

// High 32 bits of the product of two 24-bit values (v_mul_hi_u32_u24, a full-rate
// instruction; the 32-bit v_mul_hi_u32 is not).
__device__ __forceinline__ uint32_t mul_hi_u24(uint32_t a, uint32_t b) {
    uint32_t r;
    asm("v_mul_hi_u32_u24 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
    return r;
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
    const uint32_t delta = mx - mn, d = 2u * delta;
    // The reference's level is the number of c in 0..6 with 14 v >= delta (1 + 2c) -
    // (4 - 14 mn), that is y >= d (c + 1) for y = 14 (v - mn) + 4 + delta and
    // d = 2 delta: so min(floor(y / d), 7).  This takes q = floor(y / d) + 1 =
    // floor(y3 / d), y3 = y + d in [7, 4339], as (y3 m) >> 22 with m =
    // ceil(2^22 / d): y3 m / 2^22 exceeds y3 / d by less than 4339 / 2^22 = 0.00103,
    // below the gap 1 / 510 between a quotient that is no integer and the next one
    // -- checked for every delta and value on the CPU.  In 24-bit arithmetic:
    // y3 2^10 < 2^24 and m <= 2^21, so q is the high word of the 24 x 24-bit product
    // (y3 2^10) m.  m = floor((2^22 - 1) / d) + 1 from a reciprocal estimate
    // (within one of the floor) corrected by its remainder.
    uint32_t f = (uint32_t)((float)0x3FFFFF * __builtin_amdgcn_rcpf((float)d));
    int r = 0x3FFFFF - (int)(f * d);
    if (r < 0) {
        f -= 1u;
        r += (int)d;
    }
    if (r >= (int)d) f += 1u;
    const uint32_t m = f + 1u;
    const uint32_t k = (4u + 3u * delta - 14u * mn) << 10;  // y3 2^10 = 14 v 2^10 + k (mod 2^32)
    // q lies in 1..10.  The selector order {1,7,6,5,4,3,2,0} of levels 0..7 is looked
    // up by v_perm_b32 from the bytes {0, 1, 7, 6, 5, 4, 3, 2} of the two table words
    // at index q; indices 8-10 give the sign byte of a positive byte, 0, level 7's
    // selector, and the result's other bytes index byte 0, 0.
    uint32_t lo = mx | (mn << 8), hi = 0;
#pragma unroll
    for (int i = 0; i < 16; i++) {
        const uint32_t q = mul_hi_u24(__umul24(v[i], 14u << 10) + k, m);
        const uint32_t sel = __builtin_amdgcn_perm(0x02030405u, 0x06070100u, q);
        const int b = 16 + 3 * i;  // the 3-bit selector of texel i at bit b of the block
        if (b < 32) lo |= sel << b;
        if (b > 29) hi |= b >= 32 ? sel << (b - 32) : sel >> (32 - b);
    }
    return make_uint2(lo, hi);
}

// The 16 texels of block gid (index x + 4y) for one lane: four 16-byte row loads
// inside the texture, the wrap rule at its edges.
__device__ __forceinline__ void load_block16(const uint8_t *__restrict__ rgba, int w, int h, int bw, int gid,
                                             uint32_t (&px)[16]) {
    const int bx = gid % bw, by = gid / bw;
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
}

// Stores block gid: the colour block, and for BC3 the alpha block of its texels
// before it.
__device__ __forceinline__ void store_lane_block(uint8_t *__restrict__ out, uint32_t gid, const uint32_t (&px)[16],
                                                 uint2 colour, bool bc3) {
    if (bc3) {
        uint32_t a[16];
#pragma unroll
        for (int i = 0; i < 16; i++) a[i] = px[i] >> 24;
        const uint2 alpha = bc4_encode_values(a);
        reinterpret_cast<uint4 *>(out)[gid] = make_uint4(alpha.x, alpha.y, colour.x, colour.y);
    } else {
        reinterpret_cast<uint2 *>(out)[gid] = colour;
    }
}

// Job of the workgroup's unit: binary search on unit_begin.
__device__ __forceinline__ const BatchJob &unit_job(const BatchJob *__restrict__ jobs, int njobs, uint32_t u) {
    int lo = 0, hi = njobs - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (jobs[mid].unit_begin <= u) lo = mid;
        else hi = mid - 1;
    }
    return jobs[lo];
}

}  // namespace dev
}  // namespace qtx
