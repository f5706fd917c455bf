// bc1_device.h -- gfx950 device code of the BC1 encoder (also the colour half of BC3).
//
// Work decomposition: one 4x4 block per group of G = 16 lanes (four blocks per
// wave64).  The block's 16 texels sit in LDS and every lane reads them.
//   * The cheap, inherently sequential steps of BC1Encoder::EncodeBlock
//     (quicktex/s3tc/bc1/BC1Encoder.cpp:261-342) are executed redundantly by all
//     16 lanes, which keeps every group-uniform value in registers.
//   * Cluster fit (BC1Encoder.cpp:877-928) spreads its orderings over the
//     lanes; its trials are independent (each starts from the pass's input), so
//     the result is an argmin over (error, ordering index).
//   * Endpoint search (BC1Encoder.cpp:930-996) is sequential; lanes evaluate a
//     window of the next 16 rounds speculatively from the current state and the
//     first improving one is taken, which reproduces the sequential result.
//
// Selector evaluation uses v_dot4_u32_u8: for palette colour c and texel p,
// |p - c|^2 = p.p - (2 p.c - c.c), so the argmin over the palette needs one dot
// product per colour and the block error is sum(p.p) - sum(best).
//
// Bit-exactness: FP32 operations keep the reference's order, contraction is off
// (-ffp-contract=off), division is IEEE, and float->int conversions follow
// x86 cvttss2si (out of range -> INT_MIN) as the reference binary does.
#pragma once
#include <hip/hip_runtime.h>

#include <climits>
#include <cstdint>

#include "qtx_internal.h"

namespace qtx {
namespace dev {

constexpr int G = 16;  // lanes per block

__device__ __forceinline__ uint32_t chn(uint32_t c, int i) { return (c >> (8 * i)) & 0xFFu; }
__device__ __forceinline__ uint32_t rgb(uint32_t r, uint32_t g, uint32_t b) { return r | (g << 8) | (b << 16); }
__device__ __forceinline__ uint32_t dot4(uint32_t a, uint32_t b) { return __builtin_amdgcn_udot4(a, b, 0u, false); }
__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }
__device__ __forceinline__ uint32_t s5to8(uint32_t v) { return (v << 3) | (v >> 2); }
__device__ __forceinline__ uint32_t s6to8(uint32_t v) { return (v << 2) | (v >> 4); }
__device__ __forceinline__ uint32_t s8to5(uint32_t v) { uint32_t t = v * 31u + 128u; return (t + (t >> 8)) >> 8; }
__device__ __forceinline__ uint32_t s8to6(uint32_t v) { uint32_t t = v * 63u + 128u; return (t + (t >> 8)) >> 8; }
__device__ __forceinline__ bool is_black(uint32_t c) { return (chn(c, 0) | chn(c, 1) | chn(c, 2)) < 4u; }
// 565-unscaled colour (r5 | g6 << 8 | b5 << 16) -> 8-bit RGB
__device__ __forceinline__ uint32_t expand565(uint32_t c) { return rgb(s5to8(chn(c, 0)), s6to8(chn(c, 1)), s5to8(chn(c, 2))); }
__device__ __forceinline__ uint32_t to565(uint32_t c) { return rgb(s8to5(chn(c, 0)), s8to6(chn(c, 1)), s8to5(chn(c, 2))); }
__device__ __forceinline__ uint32_t pack565(uint32_t c) { return (chn(c, 0) << 11) | (chn(c, 1) << 5) | chn(c, 2); }

// x86 cvttss2si: truncation, INT_MIN for NaN and out-of-range values.
__device__ __forceinline__ int cvt_trunc(float x) {
    return (x >= -2147483648.0f && x < 2147483648.0f) ? (int)x : INT_MIN;
}

// ---------------------------------------------------------------------------
// Group (16-lane) primitives inside a wave64.
__device__ __forceinline__ int glane() { return (int)(threadIdx.x & (G - 1)); }
template <class T> __device__ __forceinline__ T gshfl(T v, int src) { return __shfl(v, src, G); }
template <class T> __device__ __forceinline__ T gxor(T v, int m) { return __shfl_xor(v, m, G); }
template <class T> __device__ __forceinline__ T gup(T v, int d) { return __shfl_up(v, (unsigned)d, G); }
__device__ __forceinline__ uint32_t gballot(bool p) {
    const uint64_t b = __ballot(p);
    return (uint32_t)(b >> (threadIdx.x & 63 & ~(G - 1))) & 0xFFFFu;
}

// ---------------------------------------------------------------------------
// Palettes: Interpolator::InterpolateBC1 (Interpolator.cpp:57-73, Nvidia :110-127),
// returned in the encoder's internal selector order: 4-colour {c0, c2, c3, c1},
// 3-colour {c0, c2, c1, black} (BC1Encoder.cpp:687-693).  Alpha bytes are 0.
template <int INTERP> __device__ __forceinline__ uint32_t third(uint32_t a, uint32_t b) {
    if (INTERP == QTX_INTERP_IDEAL_ROUND) return (2 * a + b + 1) / 3;
    if (INTERP == QTX_INTERP_AMD) return (a * 43 + b * 21 + 32) >> 6;
    return (2 * a + b) / 3;
}
template <int INTERP> __device__ __forceinline__ uint32_t half(uint32_t a, uint32_t b) {
    if (INTERP == QTX_INTERP_AMD) return (a + b + 1) >> 1;
    return (a + b) / 2;
}
__device__ __forceinline__ uint32_t nv_third5(uint32_t a, uint32_t b) { return (((2 * a + b) * 22) / 8u) & 0xFFu; }
__device__ __forceinline__ uint32_t nv_half5(uint32_t a, uint32_t b) { return (((a + b) * 33) / 8u) & 0xFFu; }
__device__ __forceinline__ uint32_t nv_third6(uint32_t a, uint32_t b) {
    const int d = (int)b - (int)a;
    return (uint32_t)((256 * (int)a + d / 4 + 128 + d * 80) >> 8) & 0xFFu;
}
__device__ __forceinline__ uint32_t nv_half6(uint32_t a, uint32_t b) {
    const int d = (int)b - (int)a;
    return (uint32_t)((256 * (int)a + d / 4 + 128 + d * 128) >> 8) & 0xFFu;
}

template <int INTERP> __device__ __forceinline__ void palette(uint32_t lo, uint32_t hi, bool three, uint32_t c[4]) {
    const uint32_t c0 = expand565(lo), c1 = expand565(hi);
    uint32_t c2, c3;
    if (INTERP == QTX_INTERP_NVIDIA) {
        const uint32_t lr = chn(lo, 0), lg = chn(lo, 1), lb = chn(lo, 2);
        const uint32_t hr = chn(hi, 0), hg = chn(hi, 1), hb = chn(hi, 2);
        if (!three) {
            c2 = rgb(nv_third5(lr, hr), nv_third6(lg, hg), nv_third5(lb, hb));
            c3 = rgb(nv_third5(hr, lr), nv_third6(hg, lg), nv_third5(hb, lb));
        } else {
            c2 = rgb(nv_half5(lr, hr), nv_half6(lg, hg), nv_half5(lb, hb));
            c3 = 0;
        }
    } else {
        const uint32_t r0 = chn(c0, 0), g0 = chn(c0, 1), b0 = chn(c0, 2);
        const uint32_t r1 = chn(c1, 0), g1 = chn(c1, 1), b1 = chn(c1, 2);
        if (!three) {
            c2 = rgb(third<INTERP>(r0, r1), third<INTERP>(g0, g1), third<INTERP>(b0, b1));
            c3 = rgb(third<INTERP>(r1, r0), third<INTERP>(g1, g0), third<INTERP>(b1, b0));
        } else {
            c2 = rgb(half<INTERP>(r0, r1), half<INTERP>(g0, g1), half<INTERP>(b0, b1));
            c3 = 0;
        }
    }
    c[0] = c0;
    c[1] = c2;
    c[2] = three ? c1 : c3;
    c[3] = three ? c3 : c1;
}


// ---------------------------------------------------------------------------
// Endpoints are held 5:6:5-packed (r << 11 | g << 5 | b), unscaled.
__device__ __forceinline__ uint32_t ep_r(uint32_t e) { return e >> 11; }
__device__ __forceinline__ uint32_t ep_g(uint32_t e) { return (e >> 5) & 63u; }
__device__ __forceinline__ uint32_t ep_b(uint32_t e) { return e & 31u; }
__device__ __forceinline__ uint32_t ep_make(uint32_t r, uint32_t g, uint32_t b) { return (r << 11) | (g << 5) | b; }
__device__ __forceinline__ uint32_t ep_unpack(uint32_t e) { return rgb(ep_r(e), ep_g(e), ep_b(e)); }
__device__ __forceinline__ uint32_t to565p(uint32_t c) { return ep_make(s8to5(chn(c, 0)), s8to6(chn(c, 1)), s8to5(chn(c, 2))); }

template <int INTERP> __device__ __forceinline__ void palette565(uint32_t lo, uint32_t hi, bool three, uint32_t c[4]) {
    palette<INTERP>(ep_unpack(lo), ep_unpack(hi), three, c);
}

// ---------------------------------------------------------------------------
// The block's 16 texels live in LDS (16 words per group, index x + 4y); every
// lane of the group reads them with broadcast loads.
extern __shared__ __align__(16) uint32_t qtx_lds[];
__device__ __forceinline__ unsigned gbase() { return threadIdx.x & ~(unsigned)(G - 1); }
__device__ __forceinline__ uint32_t tex(int i) { return qtx_lds[gbase() + (unsigned)i]; }
__device__ __forceinline__ uint4 tex4(int q) { return reinterpret_cast<const uint4 *>(qtx_lds + gbase())[q]; }
__device__ __forceinline__ uint32_t pick(const uint4 &v, int k) { return k == 0 ? v.x : (k == 1 ? v.y : (k == 2 ? v.z : v.w)); }

struct Metrics {  // ColorBlock::Metrics (quicktex/ColorBlock.h:35-42), rgb packed
    uint32_t mn, mx, avg;
    int s0, s1, s2;
    bool grey, has_black;
};

// ColorBlock::GetMetrics (quicktex/ColorBlock.h:88-117)
__device__ __forceinline__ Metrics metrics(bool ignore_black) {
    uint32_t mn0 = 255, mn1 = 255, mn2 = 255, mx0 = 0, mx1 = 0, mx2 = 0;
    int s0 = 0, s1 = 0, s2 = 0, total = 0;
    bool grey = true, has_black = false;
    for (int q = 0; q < 4; q++) {
        const uint4 t = tex4(q);
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const uint32_t v = pick(t, k);
            const bool blk = is_black(v);
            has_black |= blk;
            if (ignore_black && blk) continue;
            const uint32_t r = chn(v, 0), g = chn(v, 1), bl = chn(v, 2);
            grey &= (r == g) && (r == bl);
            mn0 = min(mn0, r); mn1 = min(mn1, g); mn2 = min(mn2, bl);
            mx0 = max(mx0, r); mx1 = max(mx1, g); mx2 = max(mx2, bl);
            s0 += (int)r; s1 += (int)g; s2 += (int)bl;
            total++;
        }
    }
    Metrics m;
    m.mn = rgb(mn0, mn1, mn2);
    m.mx = rgb(mx0, mx1, mx2);
    m.s0 = s0; m.s1 = s1; m.s2 = s2;
    m.grey = grey;
    m.has_black = has_black;
    m.avg = 0;
    if (total > 0) {
        const int h = total / 2;
        m.avg = rgb((uint32_t)((s0 + h) / total), (uint32_t)((s1 + h) / total), (uint32_t)((s2 + h) / total));
    }
    return m;
}

// Encoder state for one block: EncodeResults (BC1Encoder.h:140-147).
struct Res {
    uint32_t ep;   // low | high << 16, both 5:6:5-packed
    uint32_t sel;  // 2 bits per texel, internal order
    uint32_t err;
    int mode;  // 0 (incomplete), 3, 4, 0x13
};
__device__ __forceinline__ uint32_t res_lo(const Res &r) { return r.ep & 0xFFFFu; }
__device__ __forceinline__ uint32_t res_hi(const Res &r) { return r.ep >> 16; }
__device__ __forceinline__ uint4 pack(const Res &r) { return make_uint4(r.ep, r.sel, r.err, (uint32_t)r.mode); }
__device__ __forceinline__ Res unpack(const uint4 &v) { return Res{v.x, v.y, v.z, (int)v.w}; }

// Per-block context, passed by value: the tables, and the texels' sums of squares.
struct Ctx {
    const DeviceTables *T;
    int pp;        // sum over texels of |rgb|^2
    uint32_t ppa;  // sum over texels of |rgba|^2
};

// The encoder settings are uniform for a launch; the kernel stores them in LDS
// after the texels (words 256..) and every function reads them from there.
constexpr int kLdsTexWords = 256;
__device__ __forceinline__ const BC1Settings &S() {
    return *reinterpret_cast<const BC1Settings *>(qtx_lds + kLdsTexWords);
}

// Color::PreciseRound565 (quicktex/Color.cpp:163-183)
__device__ __forceinline__ uint32_t precise_round565(Ctx cx, float r, float g, float b) {
    int ir = clampi(cvt_trunc(r * 31.0f), 0, 31);
    int ig = clampi(cvt_trunc(g * 63.0f), 0, 63);
    int ib = clampi(cvt_trunc(b * 31.0f), 0, 31);
    ir += r > cx.T->mid5[ir];
    ig += g > cx.T->mid6[ig];
    ib += b > cx.T->mid5[ib];
    return ep_make((uint32_t)ir, (uint32_t)ig, (uint32_t)ib);
}

// FindSelectors<M> (BC1Encoder.cpp:679-779), evaluated over all 16 texels.
// The reference stops early once the running error reaches the bound; such a
// trial is never accepted by any caller (they all require a strictly lower
// error), so the full error is equivalent for every decision it feeds.
// Returns (error, selectors).
template <int INTERP>
__device__ __forceinline__ uint2 find_selectors(Ctx cx, int M, int error_mode, uint32_t lo, uint32_t hi) {
    uint32_t c[4];
    palette565<INTERP>(lo, hi, (M & 0xF) == 3, c);
    const int cc0 = (int)dot4(c[0], c[0]), cc1 = (int)dot4(c[1], c[1]);
    const int cc2 = (int)dot4(c[2], c[2]), cc3 = (int)dot4(c[3], c[3]);
    int tot = 0;
    uint32_t sel = 0;
    if (error_mode == QTX_ERROR_FULL) {
        const bool four = M != QTX_COLOR_THREE;
        for (int q = 0; q < 4; q++) {
            const uint4 t = tex4(q);
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const uint32_t p = pick(t, k);
                const int u0 = 2 * (int)dot4(p, c[0]) - cc0;
                const int u1 = 2 * (int)dot4(p, c[1]) - cc1;
                const int u2 = 2 * (int)dot4(p, c[2]) - cc2;
                const int u3 = four ? 2 * (int)dot4(p, c[3]) - cc3 : INT_MIN;
                int best = u0;
                uint32_t s = 0;
                if (u1 > best) { best = u1; s = 1; }
                if (u2 > best) { best = u2; s = 2; }
                if (u3 >= best) { best = u3; s = 3; }
                tot += best;
                sel |= s << (8 * q + 2 * k);
            }
        }
        return make_uint2((uint32_t)(cx.pp - tot), sel);
    }
    if (error_mode == QTX_ERROR_CHECK2) {
        const int a0 = (int)chn(c[3], 0) - (int)chn(c[0], 0);
        const int a1 = (int)chn(c[3], 1) - (int)chn(c[0], 1);
        const int a2 = (int)chn(c[3], 2) - (int)chn(c[0], 2);
        const uint32_t sq = (uint32_t)(a0 * a0 + a1 * a1 + a2 * a2);
        const float f = 4.0f / ((float)sq + .00000125f);
        const int c0a = (int)dot4(c[0], c[3]) - cc0;  // c0 . axis
        for (int q = 0; q < 4; q++) {
            const uint4 t = tex4(q);
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const uint32_t p = pick(t, k);
                const int d0 = (int)dot4(p, c[0]), d1 = (int)dot4(p, c[1]);
                const int d2 = (int)dot4(p, c[2]), d3 = (int)dot4(p, c[3]);
                const int da = (d3 - d0) - c0a;  // (p - c0) . axis
                const float sel_f = (float)da * f + 0.5f;
                const int s = clampi(cvt_trunc(sel_f), 1, 3);
                const int u0 = 2 * d0 - cc0, u1 = 2 * d1 - cc1, u2 = 2 * d2 - cc2, u3 = 2 * d3 - cc3;
                const int ua = s == 1 ? u0 : (s == 2 ? u1 : u2);  // candidate s-1
                const int ub = s == 1 ? u1 : (s == 2 ? u2 : u3);  // candidate s
                int bs = s, bu = ub;
                if (ua == ub) {
                    if (s == 1) bs = 0;
                } else if (ua > ub) {
                    bs = s - 1;
                    bu = ua;
                }
                tot += bu;
                sel |= (uint32_t)bs << (8 * q + 2 * k);
            }
        }
        return make_uint2((uint32_t)(cx.pp - tot), sel);
    }
    // None / Faster: axis . colour j = c3.cj - c0.cj
    const int k0 = (int)dot4(c[3], c[0]) - cc0;
    const int k1 = (int)dot4(c[3], c[1]) - (int)dot4(c[0], c[1]);
    const int k2 = (int)dot4(c[3], c[2]) - (int)dot4(c[0], c[2]);
    const int k3 = cc3 - (int)dot4(c[0], c[3]);
    const int t0 = k0 + k1, t1 = k1 + k2, t2 = k2 + k3;
    const bool faster = error_mode == QTX_ERROR_FASTER;
    for (int q = 0; q < 4; q++) {
        const uint4 t = tex4(q);
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const uint32_t p = pick(t, k);
            const int d0 = (int)dot4(p, c[0]), d3 = (int)dot4(p, c[3]);
            const int dot = 2 * (d3 - d0);
            const int level = (dot <= t0) + (dot < t1) + (dot < t2);
            const int s = 3 - level;
            if (faster) {
                const uint32_t cs = s == 0 ? c[0] : (s == 1 ? c[1] : (s == 2 ? c[2] : c[3]));
                const int ccs = s == 0 ? cc0 : (s == 1 ? cc1 : (s == 2 ? cc2 : cc3));
                tot += 2 * (int)dot4(p, cs) - ccs;
            }
            sel |= (uint32_t)s << (8 * q + 2 * k);
        }
    }
    return make_uint2(faster ? (uint32_t)(cx.pp - tot) : 0u, sel);
}

// Out-of-line copy for the once-per-block call sites.
template <int INTERP>
__device__ __noinline__ uint2 find_selectors_cold(Ctx cx, int M, int error_mode, uint32_t lo, uint32_t hi) {
    return find_selectors<INTERP>(cx, M, error_mode, lo, hi);
}

// FindEndpointsSingleColor(result, color, is_3color): match-table endpoints for
// one colour (BC1Encoder.cpp:443-456).  Returns low | high << 16; err receives the
// table error.  The reference's 3-colour tables equal the 4-colour ones
// (SingleColorTable.h: the colour-count parameter is unused).
template <int INTERP> __device__ __forceinline__ uint32_t single_color(Ctx cx, uint32_t color, uint32_t &err) {
    const uint32_t mr = cx.T->match[INTERP][0][chn(color, 0)];
    const uint32_t mg = cx.T->match[INTERP][1][chn(color, 1)];
    const uint32_t mb = cx.T->match[INTERP][0][chn(color, 2)];
    const uint32_t hi = ep_make(mr & 0xFF, mg & 0xFF, mb & 0xFF);
    const uint32_t lo = ep_make((mr >> 8) & 0xFF, (mg >> 8) & 0xFF, (mb >> 8) & 0xFF);
    err = (mr >> 16) + (mg >> 16) + (mb >> 16);
    return lo | (hi << 16);
}

// FindEndpointsSingleColor(result, pixels, color, is_3color) (BC1Encoder.cpp:458-471):
// endpoints from the match tables, error against colour 2 of the endpoints held
// on entry (alpha included), all selectors 1.
template <int INTERP>
__device__ __forceinline__ Res single_color_px(Ctx cx, uint32_t ep_in, uint32_t color, bool three) {
    uint32_t c[4];
    palette565<INTERP>(ep_in & 0xFFFFu, ep_in >> 16, three, c);
    const uint32_t qc = c[1] | 0xFF000000u;  // colour 2 of the palette, alpha 255
    uint32_t pq = 0;
    for (int q = 0; q < 4; q++) {
        const uint4 t = tex4(q);
        pq += dot4(t.x, qc) + dot4(t.y, qc) + dot4(t.z, qc) + dot4(t.w, qc);
    }
    uint32_t e;
    Res r;
    r.ep = single_color<INTERP>(cx, color, e);
    r.err = cx.ppa - 2u * pq + 16u * dot4(qc, qc);
    r.sel = 0x55555555u;
    r.mode = three ? QTX_COLOR_THREE : QTX_COLOR_FOUR;
    return r;
}

// RefineEndpointsLS(result, pixels, metrics) with 4-colour weights
// (BC1Encoder.cpp:781-825; RefineBlockLS always instantiates it <FourColor>, :860).
// Returns 0xFFFFFFFF when the system is singular, else low | high << 16.
__device__ __forceinline__ uint32_t ls_from_pixels(Ctx cx, uint32_t sel, const Metrics &m) {
    int q0 = 0, q1 = 0, q2 = 0;
    for (int q = 0; q < 4; q++) {
        const uint4 t = tex4(q);
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const uint32_t v = pick(t, k);
            const int s = (int)((sel >> (8 * q + 2 * k)) & 3u);
            q0 += (int)chn(v, 0) * s;
            q1 += (int)chn(v, 1) * s;
            q2 += (int)chn(v, 2) * s;
        }
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
    const float q00_0 = (float)q0, q00_1 = (float)q1, q00_2 = (float)q2;
    const float q10_0 = (float)(m.s0 * 3) - q00_0, q10_1 = (float)(m.s1 * 3) - q00_1, q10_2 = (float)(m.s2 * 3) - q00_2;
    const uint32_t lo = precise_round565(cx, (a0 * q00_0) + (a1 * q10_0), (a0 * q00_1) + (a1 * q10_1), (a0 * q00_2) + (a1 * q10_2));
    const uint32_t hi = precise_round565(cx, (a2 * q00_0) + (a3 * q10_0), (a2 * q00_1) + (a3 * q10_1), (a2 * q00_2) + (a3 * q10_2));
    return lo | (hi << 16);
}

// RefineBlockLS<M> (BC1Encoder.cpp:852-875); every lane of the group runs it.
template <int INTERP>
__device__ __noinline__ uint4 refine_block_ls(Ctx cx, Metrics m, int M, int error_mode, int passes, uint4 rin) {
    Res r = unpack(rin);
    for (int pass = 0; pass < passes; pass++) {
        Res t;
        const uint32_t ep = ls_from_pixels(cx, r.sel, m);
        if (ep == 0xFFFFFFFFu) {
            t = single_color_px<INTERP>(cx, r.ep, m.avg, M != QTX_COLOR_FOUR);
        } else {
            const uint2 e = find_selectors_cold<INTERP>(cx, M, error_mode, ep & 0xFFFFu, ep >> 16);
            t = Res{ep, e.y, e.x, M};
        }
        if (t.ep == r.ep) break;
        if (error_mode == QTX_ERROR_NONE || t.err < r.err) r = t;
        else break;
    }
    return pack(r);
}

// FindEndpoints (BC1Encoder.cpp:473-677); every lane of the group runs it.
// Returns low | high << 16.
__device__ __noinline__ uint32_t find_endpoints(Ctx cx, Metrics m, int endpoint_mode, bool ignore_black) {
    const uint32_t mn0 = chn(m.mn, 0), mn1 = chn(m.mn, 1), mn2 = chn(m.mn, 2);
    const uint32_t mx0 = chn(m.mx, 0), mx1 = chn(m.mx, 1), mx2 = chn(m.mx, 2);
    const uint32_t av0 = chn(m.avg, 0), av1 = chn(m.avg, 1), av2 = chn(m.avg, 2);
    if (m.grey) {
        const uint32_t fr = chn(tex(0), 0);  // texel (0,0) even when it is black (:476)
        if ((int)mx0 - (int)mn0 < 2) {
            const uint32_t e = ep_make(s8to5(fr), s8to6(fr), s8to5(fr));
            return e | (e << 16);
        }
        return ep_make(s8to5(mn0), s8to6(mn0), s8to5(mn0)) | (ep_make(s8to5(mx0), s8to6(mx0), s8to5(mx0)) << 16);
    }
    if (endpoint_mode == QTX_EP_LEAST_SQUARES) {
        // :495-552.  The reference accumulates sums_xy into uninitialised storage
        // (:510); it starts from zero here.  Both ends sit at min on the primary
        // axis (:505-506), so low == high.
        const uint32_t d0 = mx0 - mn0, d1 = mx1 - mn1, d2 = mx2 - mn2;
        const int chan0 = (d0 >= d1 && d0 >= d2) ? 0 : ((d1 >= d2 && d1 >= d0) ? 1 : 2);
        const int chan1 = (chan0 + 1) % 3, chan2 = (chan0 + 2) % 3;  // the two secondary axes, in order
        uint32_t sxx = 0, sxy1 = 0, sxy2 = 0;
        for (int i = 0; i < 16; i++) {
            const uint32_t v = tex(i);
            const uint32_t x = chn(v, chan0);
            sxx += x * x;
            sxy1 += x * chn(v, chan1);
            sxy2 += x * chn(v, chan2);
        }
        const int sums[3] = {m.s0, m.s1, m.s2};
        const uint32_t sum_x = (uint32_t)(chan0 == 0 ? sums[0] : (chan0 == 1 ? sums[1] : sums[2]));
        const uint32_t sum_y1 = (uint32_t)(chan1 == 0 ? sums[0] : (chan1 == 1 ? sums[1] : sums[2]));
        const uint32_t sum_y2 = (uint32_t)(chan2 == 0 ? sums[0] : (chan2 == 1 ? sums[1] : sums[2]));
        const float x0 = (float)chn(m.mn, chan0);
        float y1l = 0.f, y1h = 0.f, y2l = 0.f, y2h = 0.f;  // secondary-axis coordinates of low and high
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
        // inset by 1/16 of the extent and scale (:543-549), per channel
        float xl = x0, xh = x0;
        {
            const float in0 = (xh - xl) / 16.0f;
            xl = (xl + in0) / 255.0f; xh = (xh - in0) / 255.0f;
            const float in1 = (y1h - y1l) / 16.0f;
            y1l = (y1l + in1) / 255.0f; y1h = (y1h - in1) / 255.0f;
            const float in2 = (y2h - y2l) / 16.0f;
            y2l = (y2l + in2) / 255.0f; y2h = (y2h - in2) / 255.0f;
        }
        // back to r, g, b order
        const float lr = chan0 == 0 ? xl : (chan1 == 0 ? y1l : y2l);
        const float lg = chan0 == 1 ? xl : (chan1 == 1 ? y1l : y2l);
        const float lb = chan0 == 2 ? xl : (chan1 == 2 ? y1l : y2l);
        const float hr = chan0 == 0 ? xh : (chan1 == 0 ? y1h : y2h);
        const float hg = chan0 == 1 ? xh : (chan1 == 1 ? y1h : y2h);
        const float hb = chan0 == 2 ? xh : (chan1 == 2 ? y1h : y2h);
        return precise_round565(cx, lr, lg, lb) | (precise_round565(cx, hr, hg, hb) << 16);
    }
    if (endpoint_mode == QTX_EP_BOUNDING_BOX || endpoint_mode == QTX_EP_BOUNDING_BOX_INT) {
        // covariance sign of r and g against b, with the reference's operator
        // precedence (:571-573, :596-598)
        int icov_xz = 0, icov_yz = 0;
        for (int i = 0; i < 16; i++) {
            const uint32_t v = tex(i);
            const int bl = (int)chn(v, 2) - (int)av2;
            icov_xz += bl * (int)chn(v, 0) - (int)av0;
            icov_yz += bl * (int)chn(v, 1) - (int)av1;
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
            return precise_round565(cx, l0, l1, l2) | (precise_round565(cx, h0, h1, h2) << 16);
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
    // PCA, :606-674
    const float fa0 = (float)av0, fa1 = (float)av1, fa2 = (float)av2;
    // covariance = Identity + sum(diff diff^T) (upper triangle), / 255, mirrored
    float c00 = 0.f, c01 = 0.f, c02 = 0.f, c11 = 0.f, c12 = 0.f, c22 = 0.f;
    for (int i = 0; i < 16; i++) {
        const uint32_t v = tex(i);
        if (ignore_black && is_black(v)) continue;
        const float d0 = (float)chn(v, 0) - fa0, d1 = (float)chn(v, 1) - fa1, d2 = (float)chn(v, 2) - fa2;
        c00 += d0 * d0; c01 += d0 * d1; c02 += d0 * d2;
        c11 += d1 * d1; c12 += d1 * d2; c22 += d2 * d2;
    }
    // the identity's ones come first in the reference; every partial sum is an
    // exact integer below 2^24 here, so the order is immaterial
    c00 = (1.0f + c00) / 255.0f; c11 = (1.0f + c11) / 255.0f; c22 = (1.0f + c22) / 255.0f;
    c01 = c01 / 255.0f; c02 = c02 / 255.0f; c12 = c12 / 255.0f;
    const float c33 = 1.0f / 255.0f;
    const float zero = 0.0f / 255.0f;
    float d0 = (float)mx0 - (float)mn0, d1 = (float)mx1 - (float)mn1, d2 = (float)mx2 - (float)mn2, d3 = 0.0f;
    if (c02 < 0) d0 = -d0;
    if (c12 < 0) d1 = -d1;
    for (uint32_t it = 0; it < S().power_iterations; it++) {
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
    float min_dot = INFINITY, max_dot = -INFINITY;
    uint32_t pmin = tex(0), pmax = tex(0);
    for (int i = 0; i < 16; i++) {
        const uint32_t v = tex(i);
        if (ignore_black && is_black(v)) continue;
        float dot = 0.0f;
        dot += (float)chn(v, 0) * ax0;
        dot += (float)chn(v, 1) * ax1;
        dot += (float)chn(v, 2) * ax2;
        dot += 0.0f * ax3;
        if (dot > max_dot) { max_dot = dot; pmax = v; }
        if (dot < min_dot) { min_dot = dot; pmin = v; }
    }
    return to565p(pmin) | (to565p(pmax) << 16);
}

// ---------------------------------------------------------------------------
// Cluster fit, RefineBlockCF<M> (BC1Encoder.cpp:877-928), orderings spread over
// the group's lanes.  The result is identical in every lane.
template <int INTERP>
__device__ __noinline__ uint4 cluster_fit(Ctx cx, Metrics m, int M, int error_mode, uint32_t orderings, uint4 rin) {
    const Res orig = unpack(rin);
    const bool four = (M & 0xF) == 4;
    const int lane = glane();
    // histogram of the input selectors -> starting ordering (Histogram.h:70-79, OrderTable.h:91-104)
    const uint32_t lob = orig.sel & 0x55555555u, hib = (orig.sel >> 1) & 0x55555555u;
    const int n3 = __popc(lob & hib), n1 = __popc(lob & ~hib), n2 = __popc(~lob & hib);
    const int n0 = 16 - n1 - n2 - n3;
    uint32_t start;
    if (four) {
        if (n0 == 16) start = cx.T->single4[0];
        else if (n1 == 16) start = cx.T->single4[1];
        else if (n2 == 16) start = cx.T->single4[2];
        else if (n3 == 16) start = cx.T->single4[3];
        else start = cx.T->hashes4[n0 | (n1 << 4) | (n2 << 8)];
    } else {
        if (n0 == 16) start = cx.T->single3[0];
        else if (n1 == 16) start = cx.T->single3[1];
        else if (n2 == 16) start = cx.T->single3[2];
        else start = cx.T->hashes3[n0 | (n1 << 4)];
    }
    // sort texels by their projection on the endpoint axis (high - low, 8-bit);
    // the projection is an exact integer, ties broken by texel index
    const uint32_t e_hi = ep_unpack(res_hi(orig)), e_lo = ep_unpack(res_lo(orig));
    const uint32_t xh = expand565(e_hi), xl = expand565(e_lo);
    const uint32_t myp = tex(lane);
    const uint32_t mykey = ((uint32_t)(0x1000000 + ((int)dot4(myp, xh) - (int)dot4(myp, xl))) << 4) | (uint32_t)lane;
    int rank = 0;
    for (int i = 0; i < 16; i++) {
        const uint32_t v = tex(i);
        const uint32_t key = ((uint32_t)(0x1000000 + ((int)dot4(v, xh) - (int)dot4(v, xl))) << 4) | (uint32_t)i;
        rank += key < mykey;
    }
    // lane k receives the texel of rank k; an inclusive scan leaves the colour
    // total of the k+1 first sorted texels in lane k (r | g << 16, and b)
    const int base = (int)(threadIdx.x & 63 & ~(G - 1));
    uint32_t sa = (uint32_t)__builtin_amdgcn_ds_permute((base + rank) * 4, (int)(chn(myp, 0) | (chn(myp, 1) << 16)));
    uint32_t sb = (uint32_t)__builtin_amdgcn_ds_permute((base + rank) * 4, (int)chn(myp, 2));
#pragma unroll
    for (int d = 1; d < G; d <<= 1) {
        const uint32_t ta = gup(sa, d), tb = gup(sb, d);
        if (lane >= d) { sa += ta; sb += tb; }
    }
    const uint32_t total_a = gshfl(sa, G - 1), total_b = gshfl(sb, G - 1);
    const float den = four ? 3.0f : 2.0f;
    const float s16_0 = (float)(total_a & 0xFFFFu), s16_1 = (float)(total_a >> 16), s16_2 = (float)total_b;

    const uint32_t q_total = S().exhaustive ? (four ? (uint32_t)kOrders4 : (uint32_t)kOrders3) : orderings;
    const uint16_t *best_row = four ? &cx.T->best4[start * kBestOrders4] : &cx.T->best3[start * kBestOrders3];
    // best trial of this lane; ties keep the earliest ordering
    uint32_t b_err = 0xFFFFFFFFu, b_q = 0xFFFFFFFFu, b_ep = 0, b_sel = 0;
    int b_mode = 0;
    const uint32_t rounds = (q_total + G - 1) / G;
    for (uint32_t kk = 0; kk < rounds; kk++) {
        const uint32_t q = kk * G + (uint32_t)lane;
        const bool active = q < q_total;
        const uint32_t hash = active ? (S().exhaustive ? q : best_row[q]) : 0u;
        const uint32_t lv = four ? cx.T->levels4[hash] : cx.T->levels3[hash];
        // prefix sums at the ordering's cumulative levels (every lane takes part)
        const int l1 = (int)(lv & 0xFF), l2 = (int)((lv >> 8) & 0xFF), l3 = (int)((lv >> 16) & 0xFF);
        const uint32_t a1 = gshfl(sa, (l1 + 15) & 15), b1 = gshfl(sb, (l1 + 15) & 15);
        const uint32_t a2 = gshfl(sa, (l2 + 15) & 15), b2 = gshfl(sb, (l2 + 15) & 15);
        const uint32_t a3 = gshfl(sa, (l3 + 15) & 15), b3 = gshfl(sb, (l3 + 15) & 15);
        if (!active) continue;
        Res t;
        if (lv >> 24) {
            t = single_color_px<INTERP>(cx, orig.ep, m.avg, !four);
        } else {
            // RefineEndpointsLS(result, sums, matrix, hash) (:827-850)
            const float4 fm = *reinterpret_cast<const float4 *>(four ? cx.T->factors4[hash] : cx.T->factors3[hash]);
            const uint32_t qa = (l1 ? a1 : 0u) + (l2 ? a2 : 0u) + ((four && l3) ? a3 : 0u);
            const uint32_t qb = (l1 ? b1 : 0u) + (l2 ? b2 : 0u) + ((four && l3) ? b3 : 0u);
            // q10 = sums[l1] + sums[l2] (+ sums[l3]) and q00 = 3 sums[16] - q10: exact integers
            const float q10_0 = (float)(qa & 0xFFFFu), q10_1 = (float)(qa >> 16), q10_2 = (float)qb;
            const float q00_0 = (s16_0 * den) - q10_0, q00_1 = (s16_1 * den) - q10_1, q00_2 = (s16_2 * den) - q10_2;
            const uint32_t lo = precise_round565(cx, (fm.x * q00_0) + (fm.y * q10_0), (fm.x * q00_1) + (fm.y * q10_1),
                                                 (fm.x * q00_2) + (fm.y * q10_2));
            const uint32_t hi = precise_round565(cx, (fm.z * q00_0) + (fm.w * q10_0), (fm.z * q00_1) + (fm.w * q10_1),
                                                 (fm.z * q00_2) + (fm.w * q10_2));
            const uint2 e = find_selectors<INTERP>(cx, M, error_mode, lo, hi);
            t = Res{lo | (hi << 16), e.y, e.x, M};
        }
        if (t.err < b_err) {
            b_err = t.err; b_q = q; b_ep = t.ep; b_sel = t.sel; b_mode = t.mode;
        }
    }
    // group argmin over (error, ordering)
    uint32_t e = b_err, q = b_q;
#pragma unroll
    for (int msk = G / 2; msk >= 1; msk >>= 1) {
        const uint32_t oe = gxor(e, msk), oq = gxor(q, msk);
        if (oe < e || (oe == e && oq < q)) { e = oe; q = oq; }
    }
    Res r = orig;
    if (e < orig.err) {
        const int src = (int)(q % G);
        r.ep = gshfl(b_ep, src);
        r.sel = gshfl(b_sel, src);
        r.mode = gshfl(b_mode, src);
        r.err = e;
    }
    return pack(r);
}

// ---------------------------------------------------------------------------
// EndpointSearch (BC1Encoder.cpp:930-996) with a speculative window of G rounds.
// Voxel i moves one endpoint by (dr, dg, db); voxel opp(i) is its reverse.
__device__ __forceinline__ int vox_d(uint32_t table, uint32_t v) { return (int)((table >> (2 * v)) & 3u) - 1; }
// 2-bit fields per voxel, value + 1
constexpr uint32_t vox_pack(const int (&d)[16]) {
    uint32_t r = 0;
    for (int i = 0; i < 16; i++) r |= (uint32_t)(d[i] + 1) << (2 * i);
    return r;
}
constexpr int kVoxR[16] = {1, 0, 0, -1, 0, 0, 1, 1, 0, -1, -1, 0, -1, 1, 0, 0};
constexpr int kVoxG[16] = {0, 1, 0, 0, -1, 0, 1, 0, 1, -1, 0, -1, 1, -1, -1, 1};
constexpr int kVoxB[16] = {0, 0, 1, 0, 0, -1, 0, 1, 1, 0, -1, -1, 0, 0, 1, -1};
constexpr uint32_t kVoxRp = vox_pack(kVoxR), kVoxGp = vox_pack(kVoxG), kVoxBp = vox_pack(kVoxB);
// opposite voxel (the reverse move), 4-bit fields
constexpr int kVoxOppT[16] = {3, 4, 5, 0, 1, 2, 9, 10, 11, 6, 7, 8, 13, 12, 15, 14};
constexpr uint64_t opp_pack(const int (&o)[16]) {
    uint64_t r = 0;
    for (int i = 0; i < 16; i++) r |= (uint64_t)o[i] << (4 * i);
    return r;
}
constexpr uint64_t kVoxOpp = opp_pack(kVoxOppT);

template <int INTERP> __device__ __noinline__ uint4 endpoint_search(Ctx cx, uint4 rin) {
    Res r = unpack(rin);
    const int lane = glane();
    const uint32_t R = S().search_rounds;
    uint32_t prev = 0, i0 = 0;
    int forbidden = -1;
    const int mode = r.mode;
    const bool three = mode == QTX_COLOR_THREE || mode == QTX_COLOR_THREE_BLACK;
    const int emode = three ? QTX_ERROR_FULL : S().error_mode;
    const int M = three ? mode : QTX_COLOR_FOUR;
    while (i0 < R) {
        // this lane's round: the lane-th round >= i0 that is not the forbidden direction
        const uint32_t f0 = forbidden >= 0 ? i0 + (((uint32_t)forbidden - (i0 & 31u)) & 31u) : 0xFFFFFFFFu;
        uint32_t i = i0 + (uint32_t)lane;
        if (i >= f0) i += 1;
        uint32_t ip = i0 + (uint32_t)lane - 1u;  // round of the previous lane
        if (lane > 0 && ip >= f0) ip += 1;
        // evaluated by the sequential loop unless an earlier round ended it
        const bool valid = i < R && (lane == 0 || ip <= prev + 32u);
        bool improved = false;
        Res t = r;
        if (valid) {
            const uint32_t v = i & 15u;
            const uint32_t ep = (i & 16u) ? res_lo(r) : res_hi(r);
            const uint32_t er = (uint32_t)clampi((int)ep_r(ep) + vox_d(kVoxRp, v), 0, 31);
            const uint32_t eg = (uint32_t)clampi((int)ep_g(ep) + vox_d(kVoxGp, v), 0, 63);
            const uint32_t eb = (uint32_t)clampi((int)ep_b(ep) + vox_d(kVoxBp, v), 0, 31);
            const uint32_t ne = ep_make(er, eg, eb);
            t.ep = (i & 16u) ? ((r.ep & 0xFFFF0000u) | ne) : ((r.ep & 0xFFFFu) | (ne << 16));
            const uint2 e = find_selectors<INTERP>(cx, M, emode, res_lo(t), res_hi(t));
            t.err = e.x;
            t.sel = e.y;
            t.mode = M;
            improved = t.err < r.err;
        }
        const uint32_t imp = gballot(improved);
        const uint32_t vld = gballot(valid);
        if (imp) {
            const int j = __builtin_ctz(imp);
            const uint32_t ij = gshfl(i, j);
            r.ep = gshfl(t.ep, j);
            r.sel = gshfl(t.sel, j);
            r.err = gshfl(t.err, j);
            r.mode = M;
            forbidden = (int)((kVoxOpp >> (4 * (ij & 15u))) & 15u) | (int)(ij & 16u);
            prev = ij;
            i0 = ij + 1;
            continue;
        }
        // no improvement in the window: stop if it ran past the last round or the
        // loop's exit test, else continue after its last round
        const uint32_t ilast = gshfl(i, G - 1);
        if (vld != 0xFFFFu || ilast > prev + 32u) break;
        i0 = ilast + 1;
    }
    return pack(r);
}

// ---------------------------------------------------------------------------
// Block writers (BC1Encoder.cpp:345-441), 8 bytes as two little-endian words.

template <int INTERP> __device__ __forceinline__ uint2 write_solid(Ctx cx, uint32_t color) {
    uint32_t mask = 0xAA, min16, max16;
    if ((chn(color, 0) | chn(color, 1) | chn(color, 2)) == 0) {
        min16 = 0; max16 = 1; mask = 0x55;
    } else {
        uint32_t e;
        const uint32_t ep = single_color<INTERP>(cx, color, e);
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

__device__ __forceinline__ uint2 write_block(const Res &r) {
    uint32_t ep1 = res_lo(r), ep0 = res_hi(r);
    uint32_t lut;  // 4 x 2-bit entries
    if (r.mode & QTX_COLOR_FOUR) {
        lut = 1u | (3u << 2) | (2u << 4) | (0u << 6);
        if (ep1 > ep0) {
            const uint32_t t = ep1; ep1 = ep0; ep0 = t;
            lut = 0u | (2u << 2) | (3u << 4) | (1u << 6);
        } else if (ep1 == ep0) {
            if (ep1 > 0) { ep1--; lut = 0u; }
            else { ep0 = 1; ep1 = 0; lut = 0x55u; }
        }
    } else {
        lut = 1u | (2u << 2) | (0u << 4) | (3u << 6);
        if (ep1 < ep0) {
            const uint32_t t = ep1; ep1 = ep0; ep0 = t;
            lut = 0u | (2u << 2) | (1u << 4) | (3u << 6);
        }
    }
    uint32_t bits = 0;
#pragma unroll
    for (int i = 0; i < 16; i++) bits |= ((lut >> (2 * ((r.sel >> (2 * i)) & 3u))) & 3u) << (2 * i);
    return make_uint2(ep0 | (ep1 << 16), bits);
}

// ---------------------------------------------------------------------------
// BC1Encoder::EncodeBlock (BC1Encoder.cpp:261-342) for one group; the texels are
// in LDS and cx.pp / cx.ppa hold their sums of squares.
template <int INTERP> __device__ __noinline__ uint2 encode_bc1(Ctx cx) {
    const uint32_t first = tex(0);
    bool solid = true;
    for (int i = 1; i < 16; i++) solid &= tex(i) == first;  // alpha included (ColorBlock.h:80-86)
    if (solid) return write_solid<INTERP>(cx, first);

    const Metrics m = metrics(false);
    const BC1Settings &s = S();
    const int error_mode = s.error_mode;  // orderings >= 1 always makes block errors necessary

    Res orig{0u, 0u, 0xFFFFFFFFu, 0}, result{0u, 0u, 0xFFFFFFFFu, 0};
    for (int round = 0; round < s.ep_passes; round++) {
        const int ep_mode = round == 1 ? QTX_EP_BOUNDING_BOX : s.endpoint_mode;
        const Res trial_orig{find_endpoints(cx, m, ep_mode, false), 0u, 0xFFFFFFFFu, 0};
        const uint2 e = find_selectors_cold<INTERP>(cx, QTX_COLOR_FOUR, error_mode, res_lo(trial_orig), res_hi(trial_orig));
        const Res t0{trial_orig.ep, e.y, e.x, QTX_COLOR_FOUR};
        const Res t = unpack(refine_block_ls<INTERP>(cx, m, QTX_COLOR_FOUR, error_mode, s.ls_passes, pack(t0)));
        if (t.err < result.err) {
            result = t;
            orig = trial_orig;
        }
    }
    if (result.err > 0) {
        for (int it = 0; it < s.cf_passes; it++)
            result = unpack(cluster_fit<INTERP>(cx, m, QTX_COLOR_FOUR, s.error_mode, s.orderings4, pack(result)));
    }
    if (result.err > 0 && (s.color_mode & QTX_COLOR_THREE)) {
        const uint2 e = find_selectors_cold<INTERP>(cx, QTX_COLOR_THREE, QTX_ERROR_FULL, res_lo(orig), res_hi(orig));
        const Res t0{orig.ep, e.y, e.x, QTX_COLOR_THREE};
        Res t = unpack(refine_block_ls<INTERP>(cx, m, QTX_COLOR_THREE, QTX_ERROR_FULL, s.ls_passes, pack(t0)));
        if (t.err > 0) {
            for (int it = 0; it < s.cf_passes; it++)
                t = unpack(cluster_fit<INTERP>(cx, m, QTX_COLOR_THREE, QTX_ERROR_FULL, s.orderings3, pack(t)));
        }
        if (t.err < result.err) result = t;
    }
    if (result.err > 0 && s.color_mode == QTX_COLOR_THREE_BLACK && m.has_black &&
        !((chn(m.mx, 0) | chn(m.mx, 1) | chn(m.mx, 2)) < 4u)) {
        const Metrics mnb = metrics(true);
        const uint32_t ep = find_endpoints(cx, mnb, QTX_EP_PCA, true);
        const uint2 e = find_selectors_cold<INTERP>(cx, QTX_COLOR_THREE_BLACK, QTX_ERROR_FULL, ep & 0xFFFFu, ep >> 16);
        const Res t0{ep, e.y, e.x, QTX_COLOR_THREE_BLACK};
        const Res t = unpack(refine_block_ls<INTERP>(cx, mnb, QTX_COLOR_THREE_BLACK, QTX_ERROR_FULL, s.ls_passes, pack(t0)));
        if (t.err < result.err) result = t;
    }
    if (result.err > 0 && s.search_rounds > 0) result = unpack(endpoint_search<INTERP>(cx, pack(result)));
    return write_block(result);
}

}  // namespace dev
}  // namespace qtx
