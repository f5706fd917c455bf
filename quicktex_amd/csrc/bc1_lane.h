// bc1_lane.h -- gfx950 device code of the BC1 encoder with one 4x4 block per lane.
//
// For encoder settings without endpoint search (levels 0-12 and any setting
// with search_rounds == 0), a lane runs BC1Encoder::EncodeBlock
// (quicktex/s3tc/bc1/BC1Encoder.cpp:261-342) sequentially on its own block:
// the 16 texels sit in registers, so the once-per-block steps need no cross-lane
// reductions, and a wave encodes 64 blocks at a time.  Cluster fit sorts the
// texels with a compare-exchange network and keeps the 16 prefix sums of the
// sorted texels in the lane's own LDS words, so each ordering reads its three
// partial sums directly.
//
// Selector searches return the full block error.  The reference stops a search
// once its running error reaches the error it must beat; no caller ever accepts
// such a trial (all of them require a strictly lower error), so the full error
// decides the same way.  Texels are held as RGB (alpha only enters the solid-block
// test and, as a per-block constant, the single-colour error).  FP32 arithmetic
// keeps the reference's order (contraction is off); sums of integers are exact
// in any order.
#pragma once
#include <hip/hip_runtime.h>

#include <climits>
#include <cstdint>

#include "bc1_device.h"

namespace qtx {
namespace dev {
namespace lane {

constexpr int kThreads = 256;       // workgroup of the static and batch lane kernels
constexpr int kQueueThreads = 768;  // workgroup of the queued kernel: one per CU, 3 waves per SIMD

// Per-workgroup LDS of the lane kernels, allocated statically (addresses known at
// compile time fold into the LDS instructions' offsets): a head, then (with
// palette tables, PAL: the queued kernel and the drain) the tables, then the
// cluster-fit prefix sums, [17][workgroup size]: sums[k][t] = sum of thread t's
// first k sorted texels (r | g << 16, b); row 0 stays zero, so a bin boundary at
// 0 needs no test.
struct LaneHead {
    uint32_t match[2][256];  // this launch's interpolator, 5- and 6-bit
    float mid5[32], mid6[64];
    uint32_t wave_state[kQueueThreads / 64][4];  // the queued kernel's per-wave dealing state
};
// Palette tables of this launch's interpolator: entry lo | hi << bits of a 5- or
// 6-bit channel pair holds that channel of the four palette colours in the
// encoder's internal order (bytes 0-3), for 4-colour and for 3-colour blocks.
// Red and blue share the 5-bit tables.
constexpr int kPal5 = 32 * 32, kPal6 = 64 * 64;
struct LanePal {
    uint32_t four5[kPal5], four6[kPal6], three5[kPal5], three6[kPal6];
};
struct LaneLdsPal {  // 148 032 bytes: one 768-thread workgroup per CU
    LaneHead head;
    LanePal pal;
    uint2 sums[17][kQueueThreads];
};
struct LaneLdsPlain {  // 37 440 bytes: four 256-thread workgroups per CU
    LaneHead head;
    uint2 sums[17][kThreads];
};
static_assert(sizeof(LaneHead) % 16 == 0 && sizeof(LanePal) % 16 == 0, "LDS sections stay 16-byte aligned");
__shared__ LaneLdsPal lane_lds_pal;
__shared__ LaneLdsPlain lane_lds_plain;
template <bool PAL> __device__ __forceinline__ LaneHead &LL() {
    if constexpr (PAL) return lane_lds_pal.head;
    else return lane_lds_plain.head;
}
__device__ __forceinline__ LanePal &LP() { return lane_lds_pal.pal; }
// this thread's sum of its first k sorted texels
template <bool PAL> __device__ __forceinline__ uint2 &lane_sum(uint32_t k) {
    if constexpr (PAL) return lane_lds_pal.sums[k][threadIdx.x];
    else return lane_lds_plain.sums[k][threadIdx.x];
}

// Optimisation barrier on the texel registers (emits no instruction): values the
// compiler derived from the texels before it are not carried past it, which keeps
// the per-channel extractions of one phase from staying live through the next.
__device__ __forceinline__ void touch(uint32_t (&px)[16]) {
#pragma unroll
    for (int i = 0; i < 16; i++) asm volatile("" : "+v"(px[i]));
}

template <int INTERP, bool PAL> __device__ __forceinline__ void load_tables(const DeviceTables *T) {
    for (int i = (int)threadIdx.x; i < 512; i += (int)blockDim.x) LL<PAL>().match[i >> 8][i & 255] = T->match[INTERP][i >> 8][i & 255];
    if (threadIdx.x < 32) LL<PAL>().mid5[threadIdx.x] = T->mid5[threadIdx.x];
    if (threadIdx.x < 64) LL<PAL>().mid6[threadIdx.x] = T->mid6[threadIdx.x];
    lane_sum<PAL>(0) = make_uint2(0, 0);
    if (PAL) {
        // computed by palette() itself, channel by channel (its channels are
        // independent), so a lookup equals the arithmetic exactly
        for (int i = (int)threadIdx.x; i < 2 * (kPal5 + kPal6); i += (int)blockDim.x) {
            const bool three = i >= kPal5 + kPal6;
            const int j = three ? i - (kPal5 + kPal6) : i;
            const bool six = j >= kPal5;
            const uint32_t k = (uint32_t)(six ? j - kPal5 : j);
            const uint32_t lo = six ? k & 63u : k & 31u, hi = six ? k >> 6 : k >> 5;
            uint32_t c[4];
            palette<INTERP>(six ? rgb(0, lo, 0) : rgb(lo, 0, 0), six ? rgb(0, hi, 0) : rgb(hi, 0, 0), three, c);
            const uint32_t e = chn(c[0], six) | (chn(c[1], six) << 8) | (chn(c[2], six) << 16) | (chn(c[3], six) << 24);
            uint32_t *t = three ? (six ? LP().three6 : LP().three5) : (six ? LP().four6 : LP().four5);
            t[k] = e;
        }
    }
}

// The palette of endpoints ep (internal order, as palette_pair), from the LDS
// tables when PAL: three table reads and a byte transpose.
template <int INTERP, bool PAL> __device__ __forceinline__ void lane_palette(uint32_t ep, bool three, uint32_t c[4]) {
    if (!PAL) {
        palette_pair<INTERP>(ep, three, c);
        return;
    }
    const uint32_t *t5 = three ? LP().three5 : LP().four5;
    const uint32_t *t6 = three ? LP().three6 : LP().four6;
    const uint32_t r = t5[((ep >> 11) & 31u) | ((ep >> 22) & 0x3E0u)];  // low r | high r << 5
    const uint32_t g = t6[((ep >> 5) & 63u) | ((ep >> 15) & 0xFC0u)];   // low g | high g << 6
    const uint32_t b = t5[(ep & 31u) | ((ep >> 11) & 0x3E0u)];          // low b | high b << 5
    // c[k] = r.byte k | g.byte k << 8 | b.byte k << 16 (v_perm_b32: selector 0-3 picks
    // a byte of the second operand, 4-7 of the first, 12 gives zero)
    const uint32_t rg01 = __builtin_amdgcn_perm(g, r, 0x05010400u);  // r0 g0 r1 g1
    const uint32_t rg23 = __builtin_amdgcn_perm(g, r, 0x07030602u);  // r2 g2 r3 g3
    c[0] = __builtin_amdgcn_perm(b, rg01, 0x0C040100u);
    c[1] = __builtin_amdgcn_perm(b, rg01, 0x0C050302u);
    c[2] = __builtin_amdgcn_perm(b, rg23, 0x0C060100u);
    c[3] = __builtin_amdgcn_perm(b, rg23, 0x0C070302u);
}

// The same palette from the endpoints' components e (low r g b, high r g b), which
// index the tables without unpacking a pair.
template <int INTERP, bool PAL>
__device__ __forceinline__ void lane_palette_parts(const uint32_t (&e)[6], bool three, uint32_t c[4]) {
    if (!PAL) {
        palette<INTERP>(rgb(e[0], e[1], e[2]), rgb(e[3], e[4], e[5]), three, c);
        return;
    }
    const uint32_t *t5 = three ? LP().three5 : LP().four5;
    const uint32_t *t6 = three ? LP().three6 : LP().four6;
    const uint32_t r = t5[e[0] | (e[3] << 5)], g = t6[e[1] | (e[4] << 6)], b = t5[e[2] | (e[5] << 5)];
    const uint32_t rg01 = __builtin_amdgcn_perm(g, r, 0x05010400u);
    const uint32_t rg23 = __builtin_amdgcn_perm(g, r, 0x07030602u);
    c[0] = __builtin_amdgcn_perm(b, rg01, 0x0C040100u);
    c[1] = __builtin_amdgcn_perm(b, rg01, 0x0C050302u);
    c[2] = __builtin_amdgcn_perm(b, rg23, 0x0C060100u);
    c[3] = __builtin_amdgcn_perm(b, rg23, 0x0C070302u);
}

// Color::PreciseRound565 (quicktex/Color.cpp:163-183)
template <bool PAL>
__device__ __forceinline__ uint32_t prnd565(float r, float g, float b) {
    int ir = clampi(cvt_trunc(r * 31.0f), 0, 31);
    int ig = clampi(cvt_trunc(g * 63.0f), 0, 63);
    int ib = clampi(cvt_trunc(b * 31.0f), 0, 31);
    ir += r > LL<PAL>().mid5[ir];
    ig += g > LL<PAL>().mid6[ig];
    ib += b > LL<PAL>().mid5[ib];
    return ep_make((uint32_t)ir, (uint32_t)ig, (uint32_t)ib);
}

// PreciseRound565 of both endpoints of a least-squares solve.  Every component
// there is bounded far inside the int range: the solve's coefficients are an
// adjugate of integer weight sums (entries <= 144) over a non-zero integer
// determinant, times 3/255 or 2/255, so |coefficient| < 2, and the texel sums are
// <= 3 x 16 x 255, so |component| < 2^17.  The truncating conversion then needs
// no out-of-range case, and the six rounding-point reads are issued together.
// e gets the 5:6:5 components, low r g b then high r g b.
template <bool PAL>
__device__ __forceinline__ void prnd565_ls(float lr, float lg, float lb, float hr, float hg, float hb, uint32_t (&e)[6]) {
    const int ilr = clampi((int)(lr * 31.0f), 0, 31), ilg = clampi((int)(lg * 63.0f), 0, 63);
    const int ilb = clampi((int)(lb * 31.0f), 0, 31), ihr = clampi((int)(hr * 31.0f), 0, 31);
    const int ihg = clampi((int)(hg * 63.0f), 0, 63), ihb = clampi((int)(hb * 31.0f), 0, 31);
    const float mlr = LL<PAL>().mid5[ilr], mlg = LL<PAL>().mid6[ilg], mlb = LL<PAL>().mid5[ilb];
    const float mhr = LL<PAL>().mid5[ihr], mhg = LL<PAL>().mid6[ihg], mhb = LL<PAL>().mid5[ihb];
    e[0] = (uint32_t)(ilr + (lr > mlr));
    e[1] = (uint32_t)(ilg + (lg > mlg));
    e[2] = (uint32_t)(ilb + (lb > mlb));
    e[3] = (uint32_t)(ihr + (hr > mhr));
    e[4] = (uint32_t)(ihg + (hg > mhg));
    e[5] = (uint32_t)(ihb + (hb > mhb));
}
__device__ __forceinline__ uint32_t ep_pair(const uint32_t (&e)[6]) {
    return ep_make(e[0], e[1], e[2]) | (ep_make(e[3], e[4], e[5]) << 16);
}

// FindEndpointsSingleColor(result, color, is_3color) (BC1Encoder.cpp:443-456): the pair
// from the match tables (the 3-colour tables equal the 4-colour ones); err gets the
// table error.
// m: the match tables of the launch's interpolator (5-bit, 6-bit), in LDS or global.
__device__ __forceinline__ uint32_t match_pair_in(const uint32_t (*m)[256], uint32_t color, uint32_t &err) {
    const uint32_t mr = m[0][chn(color, 0)];
    const uint32_t mg = m[1][chn(color, 1)];
    const uint32_t mb = m[0][chn(color, 2)];
    err = (mr >> 16) + (mg >> 16) + (mb >> 16);
    return ep_make((mr >> 8) & 0xFF, (mg >> 8) & 0xFF, (mb >> 8) & 0xFF) |
           (ep_make(mr & 0xFF, mg & 0xFF, mb & 0xFF) << 16);
}
template <bool PAL>
__device__ __forceinline__ uint32_t match_pair(uint32_t color, uint32_t &err) {
    return match_pair_in(LL<PAL>().match, color, err);
}

// WriteBlockSolid (BC1Encoder.cpp:345-393), from match tables m
__device__ __forceinline__ uint2 write_solid_in(const uint32_t (*m)[256], uint32_t color) {
    uint32_t mask = 0xAA, min16, max16;
    if ((chn(color, 0) | chn(color, 1) | chn(color, 2)) == 0) {
        min16 = 0; max16 = 1; mask = 0x55;
    } else {
        uint32_t e;
        const uint32_t ep = match_pair_in(m, color, e);
        // the 3-colour attempt uses identical tables, so it never wins (:359-364)
        min16 = ep & 0xFFFFu;
        max16 = ep >> 16;
        if (min16 == max16) {
            if (min16 > 0) { min16--; mask = 0; }
            else { max16 = 1; min16 = 0; mask = 0x55; }
        } else if (max16 < min16) {
            const uint32_t t = min16; min16 = max16; max16 = t;
            mask = 0xFF;
        }
    }
    return make_uint2(max16 | (min16 << 16), mask * 0x01010101u);
}
template <bool PAL>
__device__ __forceinline__ uint2 write_solid(uint32_t color) {
    return write_solid_in(LL<PAL>().match, color);
}

// ColorBlock::GetMetrics (quicktex/ColorBlock.h:88-117)
__device__ __forceinline__ Metrics metrics(const uint32_t (&px)[16], bool ignore_black) {
    uint32_t mn0 = 255, mn1 = 255, mn2 = 255, mx0 = 0, mx1 = 0, mx2 = 0;
    int s0 = 0, s1 = 0, s2 = 0, total = 0;
    bool grey = true, hb = false;
#pragma unroll
    for (int i = 0; i < 16; i++) {
        const uint32_t p = px[i];
        const uint32_t r = chn(p, 0), g = chn(p, 1), b = chn(p, 2);
        const bool blk = (r | g | b) < 4u;
        hb |= blk;
        const bool inc = !(ignore_black && blk);
        grey &= !inc || (r == g && r == b);
        mn0 = inc ? min(mn0, r) : mn0; mn1 = inc ? min(mn1, g) : mn1; mn2 = inc ? min(mn2, b) : mn2;
        mx0 = inc ? max(mx0, r) : mx0; mx1 = inc ? max(mx1, g) : mx1; mx2 = inc ? max(mx2, b) : mx2;
        s0 += inc ? (int)r : 0; s1 += inc ? (int)g : 0; s2 += inc ? (int)b : 0;
        total += inc;
    }
    // The flags are taken through an optimisation barrier: left to itself the
    // compiler rewrites has_black as min(r | g | b) < 4 at its use (the
    // three-colour-black pass, after the cluster fits), which keeps the 48 channel
    // values live through the whole encode.
    uint32_t flags = (grey ? 1u : 0u) | (hb ? 2u : 0u);
    asm volatile("" : "+v"(flags));
    Metrics m;
    m.mn = rgb(mn0, mn1, mn2);
    m.mx = rgb(mx0, mx1, mx2);
    m.s0 = s0; m.s1 = s1; m.s2 = s2;
    m.grey = (flags & 1u) != 0;
    m.has_black = (flags & 2u) != 0;
    m.avg = 0;
    if (total > 0) {
        const int h = total / 2;
        m.avg = rgb((uint32_t)((s0 + h) / total), (uint32_t)((s1 + h) / total), (uint32_t)((s2 + h) / total));
    }
    return m;
}

// The metrics through an optimisation barrier (emits no instruction): what a
// phase derives from them is computed in that phase, not hoisted into an earlier
// one and carried (or spilled) in between.
__device__ __forceinline__ Metrics opaque(const Metrics &in) {
    Metrics m = in;
    uint32_t flags = (m.grey ? 1u : 0u) | (m.has_black ? 2u : 0u);
    asm volatile("" : "+v"(m.mn), "+v"(m.mx), "+v"(m.avg), "+v"(m.s0), "+v"(m.s1), "+v"(m.s2), "+v"(flags));
    m.grey = (flags & 1u) != 0;
    m.has_black = (flags & 2u) != 0;
    return m;
}

// FindEndpoints (BC1Encoder.cpp:473-677).  Returns low | high << 16.
template <bool PAL>
__device__ __forceinline__ uint32_t find_endpoints(uint32_t (&px)[16], const Metrics &m_in, int endpoint_mode,
                                                   bool ignore_black, uint32_t power_iterations) {
    const Metrics m = opaque(m_in);
    const uint32_t mn0 = chn(m.mn, 0), mn1 = chn(m.mn, 1), mn2 = chn(m.mn, 2);
    const uint32_t mx0 = chn(m.mx, 0), mx1 = chn(m.mx, 1), mx2 = chn(m.mx, 2);
    const int av0 = (int)chn(m.avg, 0), av1 = (int)chn(m.avg, 1), av2 = (int)chn(m.avg, 2);
    if (m.grey) {
        const uint32_t fr = chn(px[0], 0);  // texel (0,0) even when it is black (:476)
        if ((int)mx0 - (int)mn0 < 2) {
            const uint32_t e = ep_make(s8to5(fr), s8to6(fr), s8to5(fr));
            return e | (e << 16);
        }
        return ep_make(s8to5(mn0), s8to6(mn0), s8to5(mn0)) | (ep_make(s8to5(mx0), s8to6(mx0), s8to5(mx0)) << 16);
    }
    if (endpoint_mode == QTX_EP_LEAST_SQUARES) {
        touch(px);
        // :495-552.  The reference accumulates sums_xy into uninitialised storage
        // (:510); it starts from zero here.  Both ends sit at min on the primary
        // axis (:505-506), so low == high.
        const uint32_t e0 = mx0 - mn0, e1 = mx1 - mn1, e2 = mx2 - mn2;
        const int chan0 = (e0 >= e1 && e0 >= e2) ? 0 : ((e1 >= e2 && e1 >= e0) ? 1 : 2);
        const int chan1 = (chan0 + 1) % 3, chan2 = (chan0 + 2) % 3;
        uint32_t sxx = 0, sxy1 = 0, sxy2 = 0;
#pragma unroll
        for (int i = 0; i < 16; i++) {
            const uint32_t x = chn(px[i], chan0);
            sxx += x * x;
            sxy1 += x * chn(px[i], chan1);
            sxy2 += x * chn(px[i], chan2);
        }
        const uint32_t sum_x = (uint32_t)(chan0 == 0 ? m.s0 : (chan0 == 1 ? m.s1 : m.s2));
        const uint32_t sum_y1 = (uint32_t)(chan1 == 0 ? m.s0 : (chan1 == 1 ? m.s1 : m.s2));
        const uint32_t sum_y2 = (uint32_t)(chan2 == 0 ? m.s0 : (chan2 == 1 ? m.s1 : m.s2));
        const float x0 = (float)chn(m.mn, chan0);
        float y1l = 0.f, y1h = 0.f, y2l = 0.f, y2h = 0.f;
        const float den = (float)(16u * sxx) - (float)(sum_x * sum_x);
        if (fabsf(den) > 1e-8f) {
            const float a1 = (float)((16u * sxy1) - (sum_x * sum_y1)) / den;
            const float b1 = (float)((sxx * sum_y1) - (sxy1 * sum_x)) / den;
            y1l = b1 + (a1 * x0);
            y1h = b1 + (a1 * x0);
            const float a2 = (float)((16u * sxy2) - (sum_x * sum_y2)) / den;
            const float b2 = (float)((sxx * sum_y2) - (sxy2 * sum_x)) / den;
            y2l = b2 + (a2 * x0);
            y2h = b2 + (a2 * x0);
        }
        float xl = x0, xh = x0;
        const float in0 = (xh - xl) / 16.0f;
        xl = (xl + in0) / 255.0f;
        xh = (xh - in0) / 255.0f;
        const float in1 = (y1h - y1l) / 16.0f;
        y1l = (y1l + in1) / 255.0f;
        y1h = (y1h - in1) / 255.0f;
        const float in2 = (y2h - y2l) / 16.0f;
        y2l = (y2l + in2) / 255.0f;
        y2h = (y2h - in2) / 255.0f;
        const float lr = chan0 == 0 ? xl : (chan1 == 0 ? y1l : y2l);
        const float lg = chan0 == 1 ? xl : (chan1 == 1 ? y1l : y2l);
        const float lb = chan0 == 2 ? xl : (chan1 == 2 ? y1l : y2l);
        const float hr = chan0 == 0 ? xh : (chan1 == 0 ? y1h : y2h);
        const float hg = chan0 == 1 ? xh : (chan1 == 1 ? y1h : y2h);
        const float hb = chan0 == 2 ? xh : (chan1 == 2 ? y1h : y2h);
        return prnd565<PAL>(lr, lg, lb) | (prnd565<PAL>(hr, hg, hb) << 16);
    }
    if (endpoint_mode == QTX_EP_BOUNDING_BOX || endpoint_mode == QTX_EP_BOUNDING_BOX_INT) {
        touch(px);
        // covariance sign of r and g against b, with the reference's operator
        // precedence (:571-573, :596-598): sum of (b - avg.b) * r - avg.r
        int icov_xz = 0, icov_yz = 0;
#pragma unroll
        for (int i = 0; i < 16; i++) {
            const int bl = (int)chn(px[i], 2) - av2;
            icov_xz += bl * (int)chn(px[i], 0) - av0;
            icov_yz += bl * (int)chn(px[i], 1) - av1;
        }
        if (endpoint_mode == QTX_EP_BOUNDING_BOX) {  // :553-580
            const float bias = 8.0f / 255.0f;
            float l0 = (float)mn0 / 255.0f, h0 = (float)mx0 / 255.0f;
            float l1 = (float)mn1 / 255.0f, h1 = (float)mx1 / 255.0f;
            float l2 = (float)mn2 / 255.0f, h2 = (float)mx2 / 255.0f;
            const float i0 = (h0 - l0 - bias) / 16.0f, i1 = (h1 - l1 - bias) / 16.0f, i2 = (h2 - l2 - bias) / 16.0f;
            l0 += i0; h0 -= i0;
            l1 += i1; h1 -= i1;
            l2 += i2; h2 -= i2;
            if (icov_xz < 0) { const float t = l0; l0 = h0; h0 = t; }
            if (icov_yz < 0) { const float t = l1; l1 = h1; h1 = t; }
            return prnd565<PAL>(l0, l1, l2) | (prnd565<PAL>(h0, h1, h2) << 16);
        }
        // :581-605 (arithmetic shift of a negative inset, clamp255)
        const int in0 = ((int)(mx0 - mn0) - 8) >> 4, in1 = ((int)(mx1 - mn1) - 8) >> 4, in2 = ((int)(mx2 - mn2) - 8) >> 4;
        uint32_t l0 = (uint32_t)clampi((int)mn0 + in0, 0, 255), h0 = (uint32_t)clampi((int)mx0 - in0, 0, 255);
        uint32_t l1 = (uint32_t)clampi((int)mn1 + in1, 0, 255), h1 = (uint32_t)clampi((int)mx1 - in1, 0, 255);
        const uint32_t l2 = (uint32_t)clampi((int)mn2 + in2, 0, 255), h2 = (uint32_t)clampi((int)mx2 - in2, 0, 255);
        if (icov_xz < 0) { const uint32_t t = l0; l0 = h0; h0 = t; }
        if (icov_yz < 0) { const uint32_t t = l1; l1 = h1; h1 = t; }
        return to565p(rgb(l0, l1, l2)) | (to565p(rgb(h0, h1, h2)) << 16);
    }
    // PCA, :606-674.  Covariance = Identity + sum(diff diff^T) / 255 with exact
    // integer sums.
    touch(px);
    int c00i = 0, c11i = 0, c22i = 0, c01i = 0, c02i = 0, c12i = 0;
#pragma unroll
    for (int i = 0; i < 16; i++) {
        const uint32_t p = px[i];
        const bool inc = !(ignore_black && is_black(p));
        const int d0 = inc ? (int)chn(p, 0) - av0 : 0, d1 = inc ? (int)chn(p, 1) - av1 : 0;
        const int d2 = inc ? (int)chn(p, 2) - av2 : 0;
        c00i += d0 * d0; c11i += d1 * d1; c22i += d2 * d2;
        c01i += d0 * d1; c02i += d0 * d2; c12i += d1 * d2;
    }
    const float c00 = (1.0f + (float)c00i) / 255.0f;
    const float c11 = (1.0f + (float)c11i) / 255.0f;
    const float c22 = (1.0f + (float)c22i) / 255.0f;
    const float c01 = (float)c01i / 255.0f;
    const float c02 = (float)c02i / 255.0f;
    const float c12 = (float)c12i / 255.0f;
    const float c33 = 1.0f / 255.0f;
    const float zero = 0.0f / 255.0f;
    float d0 = (float)mx0 - (float)mn0, d1 = (float)mx1 - (float)mn1, d2 = (float)mx2 - (float)mn2, d3 = 0.0f;
    if (c02 < 0) d0 = -d0;
    if (c12 < 0) d1 = -d1;
    for (uint32_t it = 0; it < power_iterations; it++) {
        // Matrix4x4 * Vector4: row r dotted with delta, terms in index order
        float n0 = 0.0f; n0 += d0 * c00; n0 += d1 * c01; n0 += d2 * c02; n0 += d3 * zero;
        float n1 = 0.0f; n1 += d0 * c01; n1 += d1 * c11; n1 += d2 * c12; n1 += d3 * zero;
        float n2 = 0.0f; n2 += d0 * c02; n2 += d1 * c12; n2 += d2 * c22; n2 += d3 * zero;
        float n3 = 0.0f; n3 += d0 * zero; n3 += d1 * zero; n3 += d2 * zero; n3 += d3 * c33;
        d0 = n0; d1 = n1; d2 = n2; d3 = n3;
    }
    float k = 0.0f;
    if (fabsf(d0) > k) k = fabsf(d0);
    if (fabsf(d1) > k) k = fabsf(d1);
    if (fabsf(d2) > k) k = fabsf(d2);
    float ax0 = 306.0f, ax1 = 601.0f, ax2 = 117.0f, ax3 = 0.0f;
    if (k >= 2) {
        const float sc = 2048.0f / k;
        ax0 = d0 * sc; ax1 = d1 * sc; ax2 = d2 * sc; ax3 = d3 * sc;
    }
    ax0 = ax0 * 16.0f; ax1 = ax1 * 16.0f; ax2 = ax2 * 16.0f; ax3 = ax3 * 16.0f;
    // first texel with the strictly smallest / largest projection (:662-669)
    float min_dot = INFINITY, max_dot = -INFINITY;
    uint32_t lo = px[0], hi = px[0];
#pragma unroll
    for (int i = 0; i < 16; i++) {
        const uint32_t p = px[i];
        if (ignore_black && is_black(p)) continue;
        float dot = 0.0f;
        dot += (float)chn(p, 0) * ax0;
        dot += (float)chn(p, 1) * ax1;
        dot += (float)chn(p, 2) * ax2;
        dot += 0.0f * ax3;
        if (dot > max_dot) { max_dot = dot; hi = p; }
        if (dot < min_dot) { min_dot = dot; lo = p; }
    }
    return to565p(lo) | (to565p(hi) << 16);
}

// Per-block constants of the error sums: pp = sum of |rgb|^2 over the texels,
// aa = sum of (alpha - 255)^2 (the alpha term of the single-colour error).
struct Blk {
    uint32_t pp, aa;
};

// FindSelectors<M> (BC1Encoder.cpp:679-779) for endpoints ep: returns the block
// error (0 for ErrorMode::None, which computes none), sel gets the selectors
// (internal order).  |p - c|^2 = p.p - (2 p.c - c.c), with v_dot4_u32_u8.
template <int INTERP, bool PAL>
__device__ __forceinline__ uint32_t find_selectors_c(const uint32_t (&px)[16], const Blk &bk, int M, int error_mode,
                                                     const uint32_t (&c)[4], uint32_t &sel) {
    const int cc0 = (int)dot4(c[0], c[0]), cc1 = (int)dot4(c[1], c[1]);
    const int cc2 = (int)dot4(c[2], c[2]), cc3 = (int)dot4(c[3], c[3]);
    int tot = 0;
    uint32_t s_all = 0;
    if (error_mode == QTX_ERROR_NONE || error_mode == QTX_ERROR_FASTER) {
        const int k0 = (int)dot4(c[3], c[0]) - cc0;
        const int k1 = (int)dot4(c[3], c[1]) - (int)dot4(c[0], c[1]);
        const int k2 = (int)dot4(c[3], c[2]) - (int)dot4(c[0], c[2]);
        const int k3 = cc3 - (int)dot4(c[0], c[3]);
        const FasterCuts cut = faster_cuts(k0 + k1, k1 + k2, k2 + k3);
        uint32_t pc = 0;
#pragma unroll
        for (int i = 0; i < 16; i++) {
            const uint32_t p = px[i];
            const int x = (int)dot4(p, c[3]) - (int)dot4(p, c[0]);
            const bool b0 = x <= cut.f0, b1 = x <= cut.f1, b2 = x <= cut.f2;
            const uint32_t cs = b0 ? c[0] : (b1 ? c[1] : (b2 ? c[2] : c[3]));
            tot += b0 ? cc0 : (b1 ? cc1 : (b2 ? cc2 : cc3));
            pc = dot4acc(p, cs, pc);
            s_all |= (3u - (uint32_t)b0 - (uint32_t)b1 - (uint32_t)b2) << (2 * i);
        }
        sel = s_all;
        return error_mode == QTX_ERROR_NONE ? 0u : bk.pp - (2u * pc - (uint32_t)tot);
    }
    if (error_mode == QTX_ERROR_CHECK2) {
        const int a0 = (int)chn(c[3], 0) - (int)chn(c[0], 0);
        const int a1 = (int)chn(c[3], 1) - (int)chn(c[0], 1);
        const int a2 = (int)chn(c[3], 2) - (int)chn(c[0], 2);
        const uint32_t axis2 = (uint32_t)(a0 * a0 + a1 * a1 + a2 * a2);
        const Check2Cuts cut = check2_cuts(axis2, check2_f(axis2), (int)dot4(c[0], c[3]) - cc0);
        const uint32_t n0 = neg_opaque((uint32_t)cc0), n1 = neg_opaque((uint32_t)cc1);
        const uint32_t n2 = neg_opaque((uint32_t)cc2), n3 = neg_opaque((uint32_t)cc3);
#pragma unroll
        for (int i = 0; i < 16; i++) {
            const uint32_t p = px[i];
            const int d0 = (int)dot4(p, c[0]), d1 = (int)dot4(p, c[1]);
            const int d2 = (int)dot4(p, c[2]), d3 = (int)dot4(p, c[3]);
            const bool ge2 = d3 - d0 >= cut.t2, ge3 = d3 - d0 >= cut.t3;
            const int k = 1 + (int)ge2 + (int)ge3;
            const int u0 = twice_plus((uint32_t)d0, n0), u1 = twice_plus((uint32_t)d1, n1);
            const int u2 = twice_plus((uint32_t)d2, n2), u3 = twice_plus((uint32_t)d3, n3);
            const int ua = ge2 ? (ge3 ? u2 : u1) : u0;  // candidate k-1
            const int ub = ge2 ? (ge3 ? u3 : u2) : u1;  // candidate k
            uint32_t s = (uint32_t)k;
            if (ua == ub) {
                if (k == 1) s = 0;
            } else if (ua > ub) {
                s = (uint32_t)(k - 1);
            }
            tot += max(ua, ub);
            s_all |= s << (2 * i);
        }
        sel = s_all;
        return bk.pp - (uint32_t)tot;
    }
    // Full: every candidate, ties to the later one only for index 3
    const bool three = M == QTX_COLOR_THREE;
    const uint32_t n0 = neg_opaque((uint32_t)cc0), n1 = neg_opaque((uint32_t)cc1);
    const uint32_t n2 = neg_opaque((uint32_t)cc2), n3 = neg_opaque((uint32_t)cc3);
#pragma unroll
    for (int i = 0; i < 16; i++) {
        const uint32_t p = px[i];
        const int u0 = twice_plus(dot4(p, c[0]), n0), u1 = twice_plus(dot4(p, c[1]), n1);
        const int u2 = twice_plus(dot4(p, c[2]), n2), u3 = twice_plus(dot4(p, c[3]), n3);
        int bu = u0;
        uint32_t s = 0;
        if (u1 > bu) { bu = u1; s = 1; }
        if (u2 > bu) { bu = u2; s = 2; }
        if (!three && u3 >= bu) { bu = u3; s = 3; }
        tot += bu;
        s_all |= s << (2 * i);
    }
    sel = s_all;
    return bk.pp - (uint32_t)tot;
}

// The error find_selectors returns, without the selectors: what cluster-fit and
// endpoint-search trials need (the winner's selectors are computed once).
template <int INTERP, bool PAL>
__device__ __forceinline__ uint32_t trial_error_c(const uint32_t (&px)[16], const Blk &bk, int M, int error_mode,
                                                  const uint32_t (&c)[4]) {
    int tot = 0;
    if (error_mode == QTX_ERROR_FULL) {
        const uint32_t n0 = neg_opaque(dot4(c[0], c[0])), n1 = neg_opaque(dot4(c[1], c[1]));
        const uint32_t n2 = neg_opaque(dot4(c[2], c[2]));
        if (M == QTX_COLOR_THREE) {
#pragma unroll
            for (int i = 0; i < 16; i++) {
                const uint32_t p = px[i];
                tot += max(max(twice_plus(dot4(p, c[0]), n0), twice_plus(dot4(p, c[1]), n1)), twice_plus(dot4(p, c[2]), n2));
            }
        } else {
            const uint32_t n3 = neg_opaque(dot4(c[3], c[3]));
#pragma unroll
            for (int i = 0; i < 16; i++) {
                const uint32_t p = px[i];
                const int u0 = twice_plus(dot4(p, c[0]), n0), u1 = twice_plus(dot4(p, c[1]), n1);
                const int u2 = twice_plus(dot4(p, c[2]), n2), u3 = twice_plus(dot4(p, c[3]), n3);
                tot += max(max(u0, u1), max(u2, u3));
            }
        }
        return bk.pp - (uint32_t)tot;
    }
    const int cc0 = (int)dot4(c[0], c[0]), cc1 = (int)dot4(c[1], c[1]);
    const int cc2 = (int)dot4(c[2], c[2]), cc3 = (int)dot4(c[3], c[3]);
    if (error_mode == QTX_ERROR_CHECK2) {
        const int a0 = (int)chn(c[3], 0) - (int)chn(c[0], 0);
        const int a1 = (int)chn(c[3], 1) - (int)chn(c[0], 1);
        const int a2 = (int)chn(c[3], 2) - (int)chn(c[0], 2);
        const uint32_t axis2 = (uint32_t)(a0 * a0 + a1 * a1 + a2 * a2);
        const Check2Cuts cut = check2_cuts(axis2, check2_f(axis2), (int)dot4(c[0], c[3]) - cc0);
        const uint32_t n0 = neg_opaque((uint32_t)cc0), n1 = neg_opaque((uint32_t)cc1);
        const uint32_t n2 = neg_opaque((uint32_t)cc2), n3 = neg_opaque((uint32_t)cc3);
#pragma unroll
        for (int i = 0; i < 16; i++) {
            const uint32_t p = px[i];
            const int d0 = (int)dot4(p, c[0]), d1 = (int)dot4(p, c[1]);
            const int d2 = (int)dot4(p, c[2]), d3 = (int)dot4(p, c[3]);
            const bool ge2 = d3 - d0 >= cut.t2, ge3 = d3 - d0 >= cut.t3;
            const int u0 = twice_plus((uint32_t)d0, n0), u1 = twice_plus((uint32_t)d1, n1);
            const int u2 = twice_plus((uint32_t)d2, n2), u3 = twice_plus((uint32_t)d3, n3);
            const int ua = ge2 ? (ge3 ? u2 : u1) : u0;
            const int ub = ge2 ? (ge3 ? u3 : u2) : u1;
            tot += max(ua, ub);
        }
        return bk.pp - (uint32_t)tot;
    }
    if (error_mode == QTX_ERROR_FASTER) {
        const int k0 = (int)dot4(c[3], c[0]) - cc0;
        const int k1 = (int)dot4(c[3], c[1]) - (int)dot4(c[0], c[1]);
        const int k2 = (int)dot4(c[3], c[2]) - (int)dot4(c[0], c[2]);
        const int k3 = cc3 - (int)dot4(c[0], c[3]);
        const FasterCuts cut = faster_cuts(k0 + k1, k1 + k2, k2 + k3);
        uint32_t pc = 0;
#pragma unroll
        for (int i = 0; i < 16; i++) {
            const uint32_t p = px[i];
            const int x = (int)dot4(p, c[3]) - (int)dot4(p, c[0]);
            const bool b0 = x <= cut.f0, b1 = x <= cut.f1, b2 = x <= cut.f2;
            const uint32_t cs = b0 ? c[0] : (b1 ? c[1] : (b2 ? c[2] : c[3]));
            tot += b0 ? cc0 : (b1 ? cc1 : (b2 ? cc2 : cc3));
            pc = dot4acc(p, cs, pc);
        }
        return bk.pp - (2u * pc - (uint32_t)tot);
    }
    uint32_t sel;
    return find_selectors_c<INTERP, PAL>(px, bk, M, error_mode, c, sel);
}

// find_selectors and trial_error for endpoints ep (low | high << 16)
template <int INTERP, bool PAL>
__device__ __forceinline__ uint32_t find_selectors(const uint32_t (&px)[16], const Blk &bk, int M, int error_mode,
                                                   uint32_t ep, uint32_t &sel) {
    uint32_t c[4];
    lane_palette<INTERP, PAL>(ep, (M & 0xF) == 3, c);
    return find_selectors_c<INTERP, PAL>(px, bk, M, error_mode, c, sel);
}
template <int INTERP, bool PAL>
__device__ __forceinline__ uint32_t trial_error(const uint32_t (&px)[16], const Blk &bk, int M, int error_mode,
                                                uint32_t ep) {
    uint32_t c[4];
    lane_palette<INTERP, PAL>(ep, (M & 0xF) == 3, c);
    return trial_error_c<INTERP, PAL>(px, bk, M, error_mode, c);
}

// FindEndpointsSingleColor(result, pixels, color, is_3color) (BC1Encoder.cpp:458-471):
// match-table endpoints, error against colour 2 of the endpoints held on entry
// (alpha included: that colour's alpha is 255), all selectors 1.
template <int INTERP, bool PAL>
__device__ __forceinline__ Res single_color_px(const uint32_t (&px)[16], const Blk &bk, uint32_t ep_in, uint32_t color,
                                               bool three) {
    uint32_t c[4];
    lane_palette<INTERP, PAL>(ep_in, three, c);
    const uint32_t qc = c[1];
    uint32_t pq = 0;
#pragma unroll
    for (int i = 0; i < 16; i++) pq += dot4(px[i], qc);
    uint32_t e;
    Res r;
    r.ep = match_pair<PAL>(color, e);
    r.err = bk.pp - 2u * pq + 16u * dot4(qc, qc) + bk.aa;
    r.sel = 0x55555555u;
    r.mode = three ? QTX_COLOR_THREE : QTX_COLOR_FOUR;
    return r;
}

// RefineEndpointsLS(result, pixels, metrics) with 4-colour weights (BC1Encoder.cpp:781-825;
// RefineBlockLS always instantiates it <FourColor>, :860).  0xFFFFFFFF when singular.
template <bool PAL>
__device__ __forceinline__ uint32_t ls_from_pixels(const uint32_t (&px)[16], uint32_t sel, const Metrics &m) {
    uint32_t qa = 0, qb = 0;
#pragma unroll
    for (int i = 0; i < 16; i++) {
        const uint32_t s = (sel >> (2 * i)) & 3u;
        qa += (chn(px[i], 0) * s) | ((chn(px[i], 1) * s) << 16);
        qb += chn(px[i], 2) * s;
    }
    const uint32_t lob = sel & 0x55555555u, hib = (sel >> 1) & 0x55555555u;
    const int n3 = __popc(lob & hib), n1 = __popc(lob & ~hib), n2 = __popc(~lob & hib);
    const int n0 = 16 - n1 - n2 - n3;
    // sum of Weights<4>[sel] = (s^2, s(3-s), s(3-s), (3-s)^2): exact small integers
    const float m0 = (float)(n1 + 4 * n2 + 9 * n3);
    const float m1 = (float)(2 * n1 + 2 * n2);
    const float m3 = (float)(9 * n0 + 4 * n1 + n2);
    const float det = m0 * m3 - m1 * m1;
    if (fabsf(det) < 1e-8f) return 0xFFFFFFFFu;
    const float scale = (3.0f / 255.0f) / det;
    const float a0 = (m3 * 1.0f) * scale, a1 = (m1 * -1.0f) * scale, a2 = (m1 * -1.0f) * scale, a3 = (m0 * 1.0f) * scale;
    const float q0 = (float)(qa & 0xFFFFu), q1 = (float)(qa >> 16), q2 = (float)qb;
    const float r0 = (float)(m.s0 * 3) - q0, r1 = (float)(m.s1 * 3) - q1, r2 = (float)(m.s2 * 3) - q2;
    uint32_t e[6];
    prnd565_ls<PAL>((a0 * q0) + (a1 * r0), (a0 * q1) + (a1 * r1), (a0 * q2) + (a1 * r2), (a2 * q0) + (a3 * r0),
               (a2 * q1) + (a3 * r1), (a2 * q2) + (a3 * r2), e);
    return ep_pair(e);
}

// RefineBlockLS<M> (BC1Encoder.cpp:852-875)
template <int INTERP, bool PAL>
__device__ __forceinline__ Res refine_block_ls(uint32_t (&px)[16], const Blk &bk, const Metrics &m_in, int M,
                                               int error_mode, int passes, Res r) {
    const Metrics m = opaque(m_in);
    for (int pass = 0; pass < passes; pass++) {
        Res t;
        touch(px);
        const uint32_t ep = ls_from_pixels<PAL>(px, r.sel, m);
        touch(px);
        if (ep == 0xFFFFFFFFu) {
            t = single_color_px<INTERP, PAL>(px, bk, r.ep, m.avg, M != QTX_COLOR_FOUR);
        } else {
            t.ep = ep;
            t.err = find_selectors<INTERP, PAL>(px, bk, M, error_mode, ep, t.sel);
            t.mode = M;
        }
        if (t.ep == r.ep) break;
        if (error_mode == QTX_ERROR_NONE || t.err < r.err) r = t;
        else break;
    }
    return r;
}

// Compare-exchange of (key, texel) pairs; keys are distinct (they carry the texel index).
__device__ __forceinline__ void cswap(uint32_t &ka, uint32_t &kb, uint32_t &pa, uint32_t &pb) {
    const bool sw = ka > kb;
    const uint32_t k0 = sw ? kb : ka, k1 = sw ? ka : kb;
    const uint32_t p0 = sw ? pb : pa, p1 = sw ? pa : pb;
    ka = k0; kb = k1; pa = p0; pb = p1;
}

// Batcher's odd-even merge sort of 16 (key, texel) pairs, fully unrolled.
__device__ __forceinline__ void sort16(uint32_t (&k)[16], uint32_t (&p)[16]) {
#pragma unroll
    for (int s = 1; s < 16; s <<= 1) {
#pragma unroll
        for (int d = s; d >= 1; d >>= 1) {
#pragma unroll
            for (int j = d % s; j + d < 16; j += 2 * d) {
#pragma unroll
                for (int i = 0; i < d; i++) {
                    const int a = i + j, b = i + j + d;
                    if (b < 16 && a / (2 * s) == b / (2 * s)) cswap(k[a], k[b], p[a], p[b]);
                }
            }
        }
    }
}

// Cluster fit, RefineBlockCF<M> (BC1Encoder.cpp:877-928).  Trials are independent
// (each starts from orig); the first strictly lowest error wins, as in the loop.
template <int INTERP, bool PAL>
__device__ __forceinline__ Res cluster_fit(uint32_t (&px)[16], const Blk &bk, const Metrics &m_in,
                                           const DeviceTables *__restrict__ T, int M, int error_mode, uint32_t orderings,
                                           bool exhaustive, const Res orig) {
    const Metrics m = opaque(m_in);
    const bool four = (M & 0xF) == 4;
    // histogram of the input selectors -> starting ordering (Histogram.h:70-79, OrderTable.h:91-104)
    const uint32_t lob = orig.sel & 0x55555555u, hib = (orig.sel >> 1) & 0x55555555u;
    const int n3 = __popc(lob & hib), n1 = __popc(lob & ~hib), n2 = __popc(~lob & hib);
    const int n0 = 16 - n1 - n2 - n3;
    uint32_t start;
    if (four) {
        if (n0 == 16) start = T->single4[0];
        else if (n1 == 16) start = T->single4[1];
        else if (n2 == 16) start = T->single4[2];
        else if (n3 == 16) start = T->single4[3];
        else start = T->hashes4[n0 | (n1 << 4) | (n2 << 8)];
    } else {
        if (n0 == 16) start = T->single3[0];
        else if (n1 == 16) start = T->single3[1];
        else if (n2 == 16) start = T->single3[2];
        else start = T->hashes3[n0 | (n1 << 4)];
    }
    // texels sorted by their projection on high - low (8-bit), ties by index: the
    // projection is an exact integer, so the float product of the reference and
    // this dot product agree
    const uint32_t xh = expand565(ep_unpack(orig.ep >> 16)), xl = expand565(ep_unpack(orig.ep & 0xFFFFu));
    uint32_t key[16], sp[16];
#pragma unroll
    for (int i = 0; i < 16; i++) {
        const uint32_t p = px[i];
        key[i] = ((uint32_t)(0x1000000 + ((int)dot4(p, xh) - (int)dot4(p, xl))) << 4) | (uint32_t)i;
        sp[i] = p;
    }
    sort16(key, sp);
    touch(px);
    // prefix sums of the sorted texels into this thread's LDS words
    uint32_t sa = 0, sb = 0;
#pragma unroll
    for (int i = 0; i < 16; i++) {
        sa += chn(sp[i], 0) | (chn(sp[i], 1) << 16);
        sb += chn(sp[i], 2);
        lane_sum<PAL>((uint32_t)i + 1) = make_uint2(sa, sb);
    }
    const float den = four ? 3.0f : 2.0f;
    const float s16_0 = (float)(sa & 0xFFFFu), s16_1 = (float)(sa >> 16), s16_2 = (float)sb;

    touch(px);
    const uint32_t q_total = exhaustive ? (four ? (uint32_t)kOrders4 : (uint32_t)kOrders3) : orderings;
    const uint16_t *best_row = four ? &T->best4[start * kBestOrders4] : &T->best3[start * kBestOrders3];
    const uint32_t *levels = four ? T->levels4 : T->levels3;
    const float4 *factors = four ? reinterpret_cast<const float4 *>(T->factors4) : reinterpret_cast<const float4 *>(T->factors3);
    // The table reads run ahead of the trials: the ordering of trial q + 2 and the
    // levels and factors of trial q + 1 are loaded while trial q computes.  Reads
    // past the last trial fetch the last ordering again (always in bounds).
    const uint32_t q_last = q_total - 1;
    uint32_t h_next = exhaustive ? min(1u, q_last) : best_row[min(1u, q_last)];
    uint32_t h_cur = exhaustive ? 0u : best_row[0];
    uint32_t lv_next = levels[h_cur];
    float4 fm_next = factors[h_cur];
    // wait for the first trial's tables here, so that no load of the loop's entry
    // is still counted as outstanding inside the loop (where the wait would then
    // also cover the loads issued ahead)
    asm volatile("" ::"v"(lv_next), "v"(fm_next.x), "v"(fm_next.y), "v"(fm_next.z), "v"(fm_next.w));
    // trials keep (error, endpoints, single-colour flag); the winner's selectors
    // are computed once at the end
    uint32_t b_err = orig.err, b_ep = orig.ep;
    bool b_single = false, improved = false;
    // The trial of a single-bin ordering (FindEndpointsSingleColor, :458-471) depends
    // only on the block and orig.  With many orderings (the queued kernel, PAL) it
    // is computed once here and selected where such an ordering comes up, after the
    // least-squares trial that every lane runs (its factors are zero there; its
    // result is dropped), so that no lane branches off; with few, where a branch
    // is cheaper than the up-front trial, it is computed where it comes up.
    Res single{0u, 0u, 0u, 0};
    if (PAL) single = single_color_px<INTERP, PAL>(px, bk, orig.ep, m.avg, !four);
    for (uint32_t q = 0; q < q_total; q++) {
        const uint32_t lv = lv_next;
        const float4 fm = fm_next;
        lv_next = levels[h_next];
        fm_next = factors[h_next];
        const uint32_t q2 = min(q + 2u, q_last);
        h_next = exhaustive ? q2 : best_row[q2];
        // RefineEndpointsLS(result, sums, matrix, hash) (:827-850):
        // q10 = sums[l1] + sums[l2] (+ sums[l3]), q00 = d * sums[16] - q10, exact integers
        const auto ls_trial = [&](uint32_t &ep) {
            const uint32_t l1 = lv & 0xFF, l2 = (lv >> 8) & 0xFF, l3 = (lv >> 16) & 0xFF;
            const uint2 v1 = lane_sum<PAL>(l1), v2 = lane_sum<PAL>(l2), v3 = lane_sum<PAL>(four ? l3 : 0u);
            const uint32_t qa = v1.x + v2.x + v3.x, qb = v1.y + v2.y + v3.y;
            const float q10_0 = (float)(qa & 0xFFFFu), q10_1 = (float)(qa >> 16), q10_2 = (float)qb;
            const float q00_0 = (s16_0 * den) - q10_0, q00_1 = (s16_1 * den) - q10_1, q00_2 = (s16_2 * den) - q10_2;
            uint32_t e[6], c[4];
            prnd565_ls<PAL>((fm.x * q00_0) + (fm.y * q10_0), (fm.x * q00_1) + (fm.y * q10_1),
                       (fm.x * q00_2) + (fm.y * q10_2), (fm.z * q00_0) + (fm.w * q10_0),
                       (fm.z * q00_1) + (fm.w * q10_1), (fm.z * q00_2) + (fm.w * q10_2), e);
            lane_palette_parts<INTERP, PAL>(e, !four, c);
            ep = ep_pair(e);
            return trial_error_c<INTERP, PAL>(px, bk, M, error_mode, c);
        };
        const bool is_single = (lv >> 24) != 0;
        uint32_t t_err, t_ep;
        if (PAL) {
            uint32_t ls_ep;
            const uint32_t ls_err = ls_trial(ls_ep);
            t_err = is_single ? single.err : ls_err;
            t_ep = is_single ? single.ep : ls_ep;
        } else if (is_single) {
            const Res t = single_color_px<INTERP, PAL>(px, bk, orig.ep, m.avg, !four);
            t_err = t.err;
            t_ep = t.ep;
        } else {
            t_err = ls_trial(t_ep);
        }
        if (t_err < b_err) {
            b_err = t_err;
            b_ep = t_ep;
            b_single = is_single;
            improved = true;
        }
        if (t_err == 0) break;
    }
    if (!improved) return orig;
    if (b_single) return Res{b_ep, 0x55555555u, b_err, four ? QTX_COLOR_FOUR : QTX_COLOR_THREE};
    Res best{b_ep, 0u, b_err, M};
    (void)find_selectors<INTERP, PAL>(px, bk, M, error_mode, b_ep, best.sel);
    return best;
}

// Where an endpoint search stands: the next round, the last improving round, the
// forbidden move (-1: none) and whether any round has improved.
struct SearchState {
    uint32_t next, prev;
    int forbidden;
    bool improved;
};

// Rounds st.next.. of EndpointSearch (BC1Encoder.cpp:930-996), the reference's
// sequential loop: round i moves the low (i & 16) or high endpoint by voxel i & 15,
// skips the reverse of the last improving move, and stops 33 rounds after the last
// improvement.  Returns true when the search is over, false when it reached round
// `stop` first (st then holds where to resume).  The selectors are left to
// search_finish.
template <int INTERP, bool PAL>
__device__ __forceinline__ bool search_rounds(const uint32_t (&px)[16], const Blk &bk, uint32_t rounds,
                                              int four_error_mode, Res &r, SearchState &st, uint32_t stop) {
    const bool three = r.mode == QTX_COLOR_THREE || r.mode == QTX_COLOR_THREE_BLACK;
    const int M = three ? r.mode : QTX_COLOR_FOUR;
    const int emode = three ? QTX_ERROR_FULL : four_error_mode;
    for (uint32_t i = st.next; i < rounds; i++) {
        if (i >= stop) {
            st.next = i;
            return false;
        }
        if ((int)(i & 31u) == st.forbidden) continue;
        const uint32_t v = i & 15u;
        const uint32_t e = (i & 16u) ? (r.ep & 0xFFFFu) : (r.ep >> 16);
        const uint32_t ne = ep_make((uint32_t)clampi((int)ep_r(e) + vox_d(kVoxRp, v), 0, 31),
                                    (uint32_t)clampi((int)ep_g(e) + vox_d(kVoxGp, v), 0, 63),
                                    (uint32_t)clampi((int)ep_b(e) + vox_d(kVoxBp, v), 0, 31));
        const uint32_t ep = (i & 16u) ? ((r.ep & 0xFFFF0000u) | ne) : ((r.ep & 0xFFFFu) | (ne << 16));
        const uint32_t err = trial_error<INTERP, PAL>(px, bk, M, emode, ep);
        if (err < r.err) {
            r.ep = ep;
            r.err = err;
            st.improved = true;
            st.forbidden = (int)((kVoxOpp >> (4 * v)) & 15u) | (int)(i & 16u);
            st.prev = i;
        }
        if (i - st.prev > 32u) return true;
    }
    return true;
}

// The selectors of a finished search: those of its last improvement, a function
// of the endpoints and the mode (the mode of a search never changes: it is the
// block's), computed once here rather than at every improvement.
template <int INTERP, bool PAL>
__device__ __forceinline__ void search_finish(const uint32_t (&px)[16], const Blk &bk, int four_error_mode, Res &r,
                                              const SearchState &st) {
    if (!st.improved) return;
    const bool three = r.mode == QTX_COLOR_THREE || r.mode == QTX_COLOR_THREE_BLACK;
    (void)find_selectors<INTERP, PAL>(px, bk, three ? r.mode : QTX_COLOR_FOUR, three ? QTX_ERROR_FULL : four_error_mode,
                                      r.ep, r.sel);
}

// One speculative window of a search (EndpointSearch, BC1Encoder.cpp:930-996)
// run by the G lanes of a group (lanes G*k .. G*k+G-1 of the wave, each holding
// the block's texels and the same state): lane j evaluates the j-th round from
// st.next that the sequential loop would not skip.  The first improving round is
// taken, as the sequential loop takes it, and the next window starts after it;
// with no improvement the window ends the search where the sequential loop would
// (a round past the last, or one more than 32 rounds after the last improvement).
// Returns true when the search is over.  Every lane of the group must call it.
template <int INTERP, bool PAL, int G>
__device__ __forceinline__ bool search_window(const uint32_t (&px)[16], const Blk &bk, uint32_t R,
                                              int four_error_mode, Res &r, SearchState &st) {
    static_assert(G >= 2 && G <= 16 && 64 % G == 0, "group of 2..16 lanes");
    const uint32_t lane = threadIdx.x & 63u, j = lane % (uint32_t)G, gbase = lane - j;
    constexpr uint32_t kAll = (1u << G) - 1u;
    const bool three = r.mode == QTX_COLOR_THREE || r.mode == QTX_COLOR_THREE_BLACK;
    const int M = three ? r.mode : QTX_COLOR_FOUR;
    const int emode = three ? QTX_ERROR_FULL : four_error_mode;
    const uint32_t i0 = st.next;
    // the forbidden round at or after i0 (at most one in a window of <= 32 rounds)
    const uint32_t f0 = st.forbidden >= 0 ? i0 + (((uint32_t)st.forbidden - (i0 & 31u)) & 31u) : 0xFFFFFFFFu;
    uint32_t i = i0 + j;
    if (i >= f0) i += 1;
    uint32_t ip = i0 + j - 1u;  // the previous lane's round
    if (j > 0 && ip >= f0) ip += 1;
    // evaluated by the sequential loop unless an earlier round ended it
    const bool valid = i < R && (j == 0 || ip <= st.prev + 32u);
    uint32_t tep = r.ep, terr = 0xFFFFFFFFu;
    if (valid) {
        const uint32_t v = i & 15u;
        const uint32_t e = (i & 16u) ? (r.ep & 0xFFFFu) : (r.ep >> 16);
        const uint32_t ne = ep_make((uint32_t)clampi((int)ep_r(e) + vox_d(kVoxRp, v), 0, 31),
                                    (uint32_t)clampi((int)ep_g(e) + vox_d(kVoxGp, v), 0, 63),
                                    (uint32_t)clampi((int)ep_b(e) + vox_d(kVoxBp, v), 0, 31));
        tep = (i & 16u) ? ((r.ep & 0xFFFF0000u) | ne) : ((r.ep & 0xFFFFu) | (ne << 16));
        terr = trial_error<INTERP, PAL>(px, bk, M, emode, tep);
    }
    const uint32_t imp = (uint32_t)(__ballot(valid && terr < r.err) >> gbase) & kAll;
    const uint32_t vmask = (uint32_t)(__ballot(valid) >> gbase) & kAll;
    if (imp) {
        const int src = (int)(gbase + (uint32_t)__builtin_ctz(imp));
        const uint32_t ij = (uint32_t)__shfl((int)i, src), nep = (uint32_t)__shfl((int)tep, src);
        const uint32_t nerr = (uint32_t)__shfl((int)terr, src);
        r.ep = nep;
        r.err = nerr;
        st.improved = true;
        st.forbidden = (int)((kVoxOpp >> (4 * (ij & 15u))) & 15u) | (int)(ij & 16u);
        st.prev = ij;
        st.next = ij + 1u;
        return st.next >= R;
    }
    const uint32_t ilast = (uint32_t)__shfl((int)i, (int)(gbase + G - 1));
    if (vmask != kAll || ilast > st.prev + 32u) return true;
    st.next = ilast + 1u;
    return st.next >= R;
}

// A search the caller may stop after its first rounds, to resume it elsewhere.
struct SearchPark {
    uint32_t stop;  // in: the round to stop at
    bool parked;    // out: the search stopped with rounds left; r and st say where
    Res r;
    SearchState st;
};

// The settings every preset of levels 13-18 shares (BC1Encoder.cpp:74-249): Full
// error, PCA endpoints, two least-squares and two endpoint passes, no exhaustive
// cluster fit.
__host__ __device__ inline bool hq_settings(const BC1Settings &s) {
    return s.error_mode == QTX_ERROR_FULL && s.endpoint_mode == QTX_EP_PCA && s.ls_passes == 2 &&
           s.ep_passes == 2 && s.exhaustive == 0;
}

// The settings every preset of levels 0-12 shares (BC1Encoder.cpp:74-249): no
// endpoint search, one endpoint pass, one cluster-fit pass, no exhaustive fit.
__host__ __device__ inline bool light_settings(const BC1Settings &s) {
    return s.search_rounds == 0 && s.ep_passes == 1 && s.cf_passes == 1 && s.exhaustive == 0;
}

// Encoder builds: for any settings, for those of levels 13-18 (hq_settings) and
// for those of levels 0-12 (light_settings).  fixed<F> returns s with the
// settings the build F assumes as constants, so that its encoder carries no
// branches and no registers for them.
enum Fixed { kFixNone = 0, kFixHQ = 1, kFixLight = 2 };
__host__ __device__ inline int fixed_build(const BC1Settings &s) {
    return hq_settings(s) ? kFixHQ : (light_settings(s) ? kFixLight : kFixNone);
}
template <int F> __device__ __forceinline__ BC1Settings fixed(const BC1Settings &s) {
    BC1Settings t = s;
    if (F == kFixHQ) {
        t.error_mode = QTX_ERROR_FULL;
        t.endpoint_mode = QTX_EP_PCA;
        t.ls_passes = 2;
        t.ep_passes = 2;
        t.exhaustive = 0;
    } else if (F == kFixLight) {
        t.search_rounds = 0;
        t.ep_passes = 1;
        t.cf_passes = 1;
        t.exhaustive = 0;
    }
    return t;
}

// True when a batch's lane kernel may take these settings: all but the exhaustive
// cluster fit (level 19), whose 969 + 153 orderings per pass hold a lane for about
// 4 ms a block -- a small level-19 job in a large batch would set the batch's
// time -- while the 16-lane encoder deals them to 16 lanes.  (Plain level-19
// launches of 262k blocks and up take the queued lane kernel.)
__host__ __device__ constexpr bool lane_settings(int exhaustive) { return exhaustive == 0; }

// BC1Encoder::EncodeBlock (BC1Encoder.cpp:261-342) for one block (texels RGBA8,
// index x + 4y), any settings.
// With pk, the endpoint search stops at round pk->stop: if it has rounds left then,
// pk holds its state (pk->parked) and the returned block is not the block's.
template <int INTERP, bool PAL>
__device__ __forceinline__ uint2 encode_bc1(const uint32_t (&rgba)[16], const BC1Settings &s,
                                            const DeviceTables *__restrict__ T, SearchPark *pk = nullptr) {
    bool solid = true;
#pragma unroll
    for (int i = 1; i < 16; i++) solid &= rgba[i] == rgba[0];  // alpha included (ColorBlock.h:80-86)
    if (solid) return write_solid<PAL>(rgba[0]);

    uint32_t px[16];
    Blk bk{0u, 0u};
#pragma unroll
    for (int i = 0; i < 16; i++) {
        px[i] = rgba[i] & 0x00FFFFFFu;
        bk.pp += dot4(px[i], px[i]);
        const uint32_t da = 255u - (rgba[i] >> 24);
        bk.aa += da * da;
    }
    touch(px);
    const Metrics m = metrics(px, false);
    touch(px);
    const int error_mode = s.error_mode;  // orderings >= 1 always makes block errors necessary
    Res orig{0u, 0u, 0xFFFFFFFFu, 0}, result{0u, 0u, 0xFFFFFFFFu, 0};
    for (int round = 0; round < s.ep_passes; round++) {
        const int ep_mode = round == 1 ? QTX_EP_BOUNDING_BOX : s.endpoint_mode;
        const uint32_t ep = find_endpoints<PAL>(px, m, ep_mode, false, s.power_iterations);
        touch(px);
        Res t{ep, 0u, 0u, QTX_COLOR_FOUR};
        t.err = find_selectors<INTERP, PAL>(px, bk, QTX_COLOR_FOUR, error_mode, ep, t.sel);
        t = refine_block_ls<INTERP, PAL>(px, bk, m, QTX_COLOR_FOUR, error_mode, s.ls_passes, t);
        if (t.err < result.err) {
            result = t;
            orig = Res{ep, 0u, 0xFFFFFFFFu, 0};
        }
    }
    touch(px);
    if (result.err > 0) {
        for (int it = 0; it < s.cf_passes; it++)
            result = cluster_fit<INTERP, PAL>(px, bk, m, T, QTX_COLOR_FOUR, s.error_mode, s.orderings4, s.exhaustive, result);
    }
    touch(px);
    if (result.err > 0 && (s.color_mode & QTX_COLOR_THREE)) {
        Res t{orig.ep, 0u, 0u, QTX_COLOR_THREE};
        t.err = find_selectors<INTERP, PAL>(px, bk, QTX_COLOR_THREE, QTX_ERROR_FULL, orig.ep, t.sel);
        t = refine_block_ls<INTERP, PAL>(px, bk, m, QTX_COLOR_THREE, QTX_ERROR_FULL, s.ls_passes, t);
        if (t.err > 0) {
            for (int it = 0; it < s.cf_passes; it++)
                t = cluster_fit<INTERP, PAL>(px, bk, m, T, QTX_COLOR_THREE, QTX_ERROR_FULL, s.orderings3, s.exhaustive, t);
        }
        if (t.err < result.err) result = t;
    }
    touch(px);
    if (result.err > 0 && s.color_mode == QTX_COLOR_THREE_BLACK && m.has_black &&
        !((chn(m.mx, 0) | chn(m.mx, 1) | chn(m.mx, 2)) < 4u)) {
        const Metrics mnb = metrics(px, true);
        touch(px);
        const uint32_t ep = find_endpoints<PAL>(px, mnb, QTX_EP_PCA, true, s.power_iterations);
        touch(px);
        Res t{ep, 0u, 0u, QTX_COLOR_THREE_BLACK};
        t.err = find_selectors<INTERP, PAL>(px, bk, QTX_COLOR_THREE_BLACK, QTX_ERROR_FULL, ep, t.sel);
        t = refine_block_ls<INTERP, PAL>(px, bk, mnb, QTX_COLOR_THREE_BLACK, QTX_ERROR_FULL, s.ls_passes, t);
        if (t.err < result.err) result = t;
    }
    if (result.err > 0 && s.search_rounds > 0) {
        touch(px);
        SearchState st{0u, 0u, -1, false};
        const bool over = search_rounds<INTERP, PAL>(px, bk, s.search_rounds, s.error_mode, result, st,
                                                     pk ? pk->stop : s.search_rounds);
        if (pk && !over) {
            pk->parked = true;
            pk->r = result;
            pk->st = st;
            return make_uint2(0u, 0u);
        }
        search_finish<INTERP, PAL>(px, bk, s.error_mode, result, st);
    }
    return write_block(result);
}

}  // namespace lane
}  // namespace dev
}  // namespace qtx
