// bc1_device.h -- gfx950 device code of the BC1 encoder (also the colour half of BC3).
//
// Work decomposition: one 4x4 block per group of G = 16 lanes (four blocks per
// wave64).  Lane i of a group owns texel i (index x + 4y); the block's texels
// also sit in LDS for the loops that need all of them in one lane.
//   * The once-per-block steps of BC1Encoder::EncodeBlock
//     (quicktex/s3tc/bc1/BC1Encoder.cpp:261-342) -- metrics, endpoint finding,
//     the first selector searches, least-squares refinement -- run
//     texel-parallel: each lane handles its texel and 16-lane DPP reductions
//     combine them.
//   * Cluster fit (BC1Encoder.cpp:877-928) deals its orderings to the lanes; its
//     trials are independent (each starts from the pass's input), so the result
//     is an argmin over (error, ordering index).
//   * Endpoint search (BC1Encoder.cpp:930-996) is sequential; lanes evaluate a
//     window of the next 16 rounds from the current state and the first
//     improving round is taken, which reproduces the sequential result.
//   Trials only need their error, so they run an error-only selector search;
//   the winner's selectors are recomputed texel-parallel.
//
// Selector search uses v_dot4_u32_u8: for palette colour c and texel p,
// |p - c|^2 = p.p - (2 p.c - c.c), so each palette colour costs one dot product
// and the block error is sum(p.p) - sum(best).
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
__device__ __forceinline__ uint32_t dot4acc(uint32_t a, uint32_t b, uint32_t acc) {
    return __builtin_amdgcn_udot4(a, b, acc, false);
}
__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }
__device__ __forceinline__ uint32_t s5to8(uint32_t v) { return (v << 3) | (v >> 2); }
__device__ __forceinline__ uint32_t s6to8(uint32_t v) { return (v << 2) | (v >> 4); }
__device__ __forceinline__ uint32_t s8to5(uint32_t v) { uint32_t t = v * 31u + 128u; return (t + (t >> 8)) >> 8; }
__device__ __forceinline__ uint32_t s8to6(uint32_t v) { uint32_t t = v * 63u + 128u; return (t + (t >> 8)) >> 8; }
__device__ __forceinline__ bool is_black(uint32_t c) { return (chn(c, 0) | chn(c, 1) | chn(c, 2)) < 4u; }
__device__ __forceinline__ uint32_t expand565(uint32_t c) { return rgb(s5to8(chn(c, 0)), s6to8(chn(c, 1)), s5to8(chn(c, 2))); }

// x86 cvttss2si: truncation, INT_MIN for NaN and out-of-range values.
__device__ __forceinline__ int cvt_trunc(float x) {
    return (x >= -2147483648.0f && x < 2147483648.0f) ? (int)x : INT_MIN;
}

// ---------------------------------------------------------------------------
// 16-lane group primitives inside a wave64.  Reductions use DPP within a row of
// 16 lanes (quad_perm, row_half_mirror, row_mirror); every lane of the group gets
// the result.
__device__ __forceinline__ int glane() { return (int)(threadIdx.x & (G - 1)); }
template <class T> __device__ __forceinline__ T gshfl(T v, int src) { return __shfl(v, src, G); }
__device__ __forceinline__ uint32_t gballot(bool p) {
    const uint64_t b = __ballot(p);
    return (uint32_t)(b >> (threadIdx.x & 63 & ~(G - 1))) & 0xFFFFu;
}
template <int CTRL> __device__ __forceinline__ uint32_t dpp(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, CTRL, 0xF, 0xF, false);
}
constexpr int kQuadSwap1 = 0xB1, kQuadSwap2 = 0x4E, kRowHalfMirror = 0x141, kRowMirror = 0x140;
__device__ __forceinline__ uint32_t gsum(uint32_t v) {
    v += dpp<kQuadSwap1>(v);
    v += dpp<kQuadSwap2>(v);
    v += dpp<kRowHalfMirror>(v);
    v += dpp<kRowMirror>(v);
    return v;
}
__device__ __forceinline__ uint32_t gor(uint32_t v) {
    v |= dpp<kQuadSwap1>(v);
    v |= dpp<kQuadSwap2>(v);
    v |= dpp<kRowHalfMirror>(v);
    v |= dpp<kRowMirror>(v);
    return v;
}
__device__ __forceinline__ uint32_t gmin(uint32_t v) {
    v = min(v, dpp<kQuadSwap1>(v));
    v = min(v, dpp<kQuadSwap2>(v));
    v = min(v, dpp<kRowHalfMirror>(v));
    v = min(v, dpp<kRowMirror>(v));
    return v;
}
__device__ __forceinline__ uint32_t gmax(uint32_t v) {
    v = max(v, dpp<kQuadSwap1>(v));
    v = max(v, dpp<kQuadSwap2>(v));
    v = max(v, dpp<kRowHalfMirror>(v));
    v = max(v, dpp<kRowMirror>(v));
    return v;
}
// inclusive prefix sum over the group's lanes (row_shr:n; lanes below n add 0)
__device__ __forceinline__ uint32_t gscan(uint32_t v) {
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xF, 0xF, false);
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xF, 0xF, false);
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xF, 0xF, false);
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xF, 0xF, false);
    return v;
}
// (key, index) with the smallest key; ties go to the smallest index.
template <int CTRL> __device__ __forceinline__ void argmin_step(uint32_t &k, uint32_t &i) {
    const uint32_t ok = dpp<CTRL>(k), oi = dpp<CTRL>(i);
    const bool take = ok < k || (ok == k && oi < i);
    k = take ? ok : k;
    i = take ? oi : i;
}
__device__ __forceinline__ void gargmin(uint32_t &k, uint32_t &i) {
    argmin_step<kQuadSwap1>(k, i);
    argmin_step<kQuadSwap2>(k, i);
    argmin_step<kRowHalfMirror>(k, i);
    argmin_step<kRowMirror>(k, i);
}
// order-preserving map of a float to an unsigned key (-0 and +0 tie, as they compare equal)
__device__ __forceinline__ uint32_t fkey(float f) {
    const uint32_t u = __float_as_uint(f == 0.0f ? 0.0f : f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
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

// lo, hi: unscaled 5:6:5 colours as bytes (r5 | g6 << 8 | b5 << 16)
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
// Endpoints are held 5:6:5-packed (r << 11 | g << 5 | b), unscaled; a pair as low | high << 16.
__device__ __forceinline__ uint32_t ep_r(uint32_t e) { return e >> 11; }
__device__ __forceinline__ uint32_t ep_g(uint32_t e) { return (e >> 5) & 63u; }
__device__ __forceinline__ uint32_t ep_b(uint32_t e) { return e & 31u; }
__device__ __forceinline__ uint32_t ep_make(uint32_t r, uint32_t g, uint32_t b) { return (r << 11) | (g << 5) | b; }
__device__ __forceinline__ uint32_t ep_unpack(uint32_t e) { return rgb(ep_r(e), ep_g(e), ep_b(e)); }
__device__ __forceinline__ uint32_t to565p(uint32_t c) { return ep_make(s8to5(chn(c, 0)), s8to6(chn(c, 1)), s8to5(chn(c, 2))); }
template <int INTERP> __device__ __forceinline__ void palette_pair(uint32_t ep, bool three, uint32_t c[4]) {
    palette<INTERP>(ep_unpack(ep & 0xFFFFu), ep_unpack(ep >> 16), three, c);
}

// ---------------------------------------------------------------------------
// Per-workgroup LDS: tables copied from DeviceTables once, then the texels of
// the workgroup's 16 blocks (16 words per group) and one settings slot per wave
// (a wave's four blocks always share their settings; waves of one workgroup may
// work on differently configured textures in a batch).
constexpr int kWaves = 4;  // waves per 256-lane workgroup
struct Lds {
    float4 f4[kOrders4];  // LS factors per ordering
    float4 f3[kOrders3];
    uint32_t tex[256];
    uint32_t lv4[kOrders4];  // cumulative levels | single-colour flag << 24
    uint32_t lv3[kOrders3];
    float mid5[32], mid6[64];
    uint32_t match[2][256];  // this launch's interpolator, 5- and 6-bit
    BC1Settings s[kWaves];
};
extern __shared__ __align__(16) uint8_t qtx_lds_raw[];
__device__ __forceinline__ Lds &L() { return *reinterpret_cast<Lds *>(qtx_lds_raw); }
__device__ __forceinline__ unsigned gbase() { return threadIdx.x & ~(unsigned)(G - 1); }
__device__ __forceinline__ uint32_t tex(int i) { return L().tex[gbase() + (unsigned)i]; }
__device__ __forceinline__ uint4 tex4(int q) { return reinterpret_cast<const uint4 *>(&L().tex[gbase()])[q]; }
__device__ __forceinline__ uint32_t pick(const uint4 &v, int k) { return k == 0 ? v.x : (k == 1 ? v.y : (k == 2 ? v.z : v.w)); }
__device__ __forceinline__ int wave_slot() { return __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6)); }
__device__ __forceinline__ const BC1Settings &S() { return L().s[wave_slot()]; }

struct Ctx {
    const DeviceTables *T;
    uint32_t px;   // this lane's texel (RGBA8)
    int pp;        // sum over the block's texels of |rgb|^2
    uint32_t ppa;  // sum over the block's texels of |rgba|^2
};

struct Metrics {  // ColorBlock::Metrics (quicktex/ColorBlock.h:35-42), rgb packed
    uint32_t mn, mx, avg;
    int s0, s1, s2;
    bool grey, has_black;
};

// Encoder state for one block: EncodeResults (BC1Encoder.h:140-147).
struct Res {
    uint32_t ep;   // low | high << 16
    uint32_t sel;  // 2 bits per texel, internal order
    uint32_t err;
    int mode;  // 0 (incomplete), 3, 4, 0x13
};

// ColorBlock::GetMetrics (quicktex/ColorBlock.h:88-117), texel-parallel
__device__ __forceinline__ Metrics metrics(const Ctx &cx, bool ignore_black) {
    const uint32_t p = cx.px;
    const bool blk = is_black(p);
    const bool inc = !(ignore_black && blk);
    const uint32_t r = chn(p, 0), g = chn(p, 1), b = chn(p, 2);
    Metrics m;
    m.has_black = gballot(blk) != 0;
    m.grey = gballot(inc && !(r == g && r == b)) == 0;
    m.mn = rgb(gmin(inc ? r : 255u), gmin(inc ? g : 255u), gmin(inc ? b : 255u));
    m.mx = rgb(gmax(inc ? r : 0u), gmax(inc ? g : 0u), gmax(inc ? b : 0u));
    const uint32_t sa = gsum(inc ? (r | (g << 16)) : 0u), sb = gsum(inc ? b : 0u);
    m.s0 = (int)(sa & 0xFFFFu);
    m.s1 = (int)(sa >> 16);
    m.s2 = (int)sb;
    const int total = __popc(gballot(inc));
    m.avg = 0;
    if (total > 0) {
        const int h = total / 2;
        m.avg = rgb((uint32_t)((m.s0 + h) / total), (uint32_t)((m.s1 + h) / total), (uint32_t)((m.s2 + h) / total));
    }
    return m;
}

// Color::PreciseRound565 (quicktex/Color.cpp:163-183)
__device__ __forceinline__ uint32_t precise_round565(float r, float g, float b) {
    int ir = clampi(cvt_trunc(r * 31.0f), 0, 31);
    int ig = clampi(cvt_trunc(g * 63.0f), 0, 63);
    int ib = clampi(cvt_trunc(b * 31.0f), 0, 31);
    ir += r > L().mid5[ir];
    ig += g > L().mid6[ig];
    ib += b > L().mid5[ib];
    return ep_make((uint32_t)ir, (uint32_t)ig, (uint32_t)ib);
}

// PreciseRound565 of both endpoints of a least-squares solve (low | high << 16).
// Every component there is bounded far inside the int range (coefficients: an
// adjugate of integer weight sums over a non-zero integer determinant, times 3/255
// or 2/255, so below 2 in magnitude; texel sums at most 3 x 16 x 255), so the
// truncating conversion needs no out-of-range case, and the six rounding-point
// reads go out together.
__device__ __forceinline__ uint32_t precise_round565_ls(float lr, float lg, float lb, float hr, float hg, float hb) {
    const int ilr = clampi((int)(lr * 31.0f), 0, 31), ilg = clampi((int)(lg * 63.0f), 0, 63);
    const int ilb = clampi((int)(lb * 31.0f), 0, 31), ihr = clampi((int)(hr * 31.0f), 0, 31);
    const int ihg = clampi((int)(hg * 63.0f), 0, 63), ihb = clampi((int)(hb * 31.0f), 0, 31);
    const float mlr = L().mid5[ilr], mlg = L().mid6[ilg], mlb = L().mid5[ilb];
    const float mhr = L().mid5[ihr], mhg = L().mid6[ihg], mhb = L().mid5[ihb];
    const uint32_t lo = ep_make((uint32_t)(ilr + (lr > mlr)), (uint32_t)(ilg + (lg > mlg)), (uint32_t)(ilb + (lb > mlb)));
    const uint32_t hi = ep_make((uint32_t)(ihr + (hr > mhr)), (uint32_t)(ihg + (hg > mhg)), (uint32_t)(ihb + (hb > mhb)));
    return lo | (hi << 16);
}

// ---------------------------------------------------------------------------
// Selector search, FindSelectors<M> (BC1Encoder.cpp:679-779).  The reference stops
// early once the running error reaches the bound; no caller ever accepts such a
// trial (they all require a strictly lower error), so the full error decides the
// same way.

// -v, opaque to the optimiser: 2 p.c - c.c is then written (p.c << 1) + (-c.c), one
// v_lshl_add_u32, where the folded form 2 p.c - c.c costs a shift and a subtract.
__device__ __forceinline__ uint32_t neg_opaque(uint32_t v) {
    uint32_t n = 0u - v;
    asm("" : "+v"(n));
    return n;
}
__device__ __forceinline__ int twice_plus(uint32_t d, uint32_t n) { return (int)((d << 1) + n); }

// The Faster (and None) selector of a texel (BC1Encoder.cpp:695-720): s = 3 - [dot <=
// t0] - [dot < t1] - [dot < t2] for dot = 2 x, x = (d3 - d0) the texel's projection on
// the axis and t_j the sums of neighbouring palette projections.  As cuts on x:
// [2x <= t] = [x <= t >> 1] and [2x < t] = [x <= (t - 1) >> 1] (arithmetic shifts round
// down).  s counts the cuts below x, which sorting them does not change, and with them
// sorted (f0 <= f1 <= f2) s = 0 up to f0, 1 up to f1, 2 up to f2 and 3 above: the
// palette entry follows from three compares without forming s.
struct FasterCuts {
    int f0, f1, f2;
};
__device__ __forceinline__ FasterCuts faster_cuts(int t0, int t1, int t2) {
    const int e0 = t0 >> 1, e1 = (t1 - 1) >> 1, e2 = (t2 - 1) >> 1;
    const int lo = min(e0, e1), hi = max(e0, e1);
    return FasterCuts{min(lo, e2), max(lo, min(hi, e2)), max(hi, e2)};
}

// ErrorMode::Check2's scale of the projection on the axis (BC1Encoder.cpp:721-749):
// 4 / (|axis|^2 + 1.25e-6), positive; the projection's candidate index follows from
// it through check2_cuts.
__device__ __forceinline__ float check2_f(uint32_t axis2) { return 4.0f / ((float)axis2 + .00000125f); }

// Check2's candidate index k = clamp(int(x * f + 0.5f), 1, 3) for the projection x =
// (p - c0) . axis, as two cuts on x: k = 1 + (x >= t2) + (x >= t3).  x is an integer
// of magnitude at most 3 x 255^2, exact as a float, and x * f + 0.5f, rounded at each
// step, does not decrease as x grows (f > 0): k reaches 2 exactly from the least x
// whose value reaches 2.0 (t2), and 3 from the least reaching 3.0 (t3).  For every
// |axis|^2 = A in [0, 3 x 255^2] those are ceil(3A / 8) or the next integer, and
// ceil(5A / 8) or the next (found by evaluating every A), so one evaluation of the
// rounded expression at the first decides each.  Returned as cuts on d3 - d0 =
// x + c0a, the form the selector loops hold.
struct Check2Cuts {
    int t2, t3;
};
__device__ __forceinline__ Check2Cuts check2_cuts(uint32_t axis2, float f, int c0a) {
    const int c2 = (int)((3u * axis2 + 7u) >> 3), c3 = (int)((5u * axis2 + 7u) >> 3);
    const int t2 = (float)c2 * f + 0.5f >= 2.0f ? c2 : c2 + 1;
    const int t3 = (float)c3 * f + 0.5f >= 3.0f ? c3 : c3 + 1;
    return Check2Cuts{t2 + c0a, t3 + c0a};
}

// Texel-parallel, with selectors: returns (error, selectors).
template <int INTERP>
__device__ __forceinline__ uint2 eval_px(const Ctx &cx, int M, int error_mode, uint32_t ep) {
    uint32_t c[4];
    palette_pair<INTERP>(ep, (M & 0xF) == 3, c);
    const uint32_t p = cx.px;
    const int d0 = (int)dot4(p, c[0]), d1 = (int)dot4(p, c[1]), d2 = (int)dot4(p, c[2]), d3 = (int)dot4(p, c[3]);
    const int u0 = 2 * d0 - (int)dot4(c[0], c[0]), u1 = 2 * d1 - (int)dot4(c[1], c[1]);
    const int u2 = 2 * d2 - (int)dot4(c[2], c[2]), u3 = 2 * d3 - (int)dot4(c[3], c[3]);
    int bu;
    uint32_t s;
    if (error_mode == QTX_ERROR_FULL) {
        bu = u0;
        s = 0;
        if (u1 > bu) { bu = u1; s = 1; }
        if (u2 > bu) { bu = u2; s = 2; }
        if (M != QTX_COLOR_THREE && u3 >= bu) { bu = u3; s = 3; }
    } else if (error_mode == QTX_ERROR_CHECK2) {
        const int a0 = (int)chn(c[3], 0) - (int)chn(c[0], 0);
        const int a1 = (int)chn(c[3], 1) - (int)chn(c[0], 1);
        const int a2 = (int)chn(c[3], 2) - (int)chn(c[0], 2);
        const uint32_t axis2 = (uint32_t)(a0 * a0 + a1 * a1 + a2 * a2);
        const Check2Cuts cut = check2_cuts(axis2, check2_f(axis2), (int)dot4(c[0], c[3]) - (int)dot4(c[0], c[0]));
        const bool ge2 = d3 - d0 >= cut.t2, ge3 = d3 - d0 >= cut.t3;
        const int k = 1 + (int)ge2 + (int)ge3;
        const int ua = ge2 ? (ge3 ? u2 : u1) : u0;  // candidate k-1
        const int ub = ge2 ? (ge3 ? u3 : u2) : u1;  // candidate k
        s = (uint32_t)k;
        bu = ub;
        if (ua == ub) {
            if (k == 1) s = 0;
        } else if (ua > ub) {
            s = (uint32_t)(k - 1);
            bu = ua;
        }
    } else {  // None / Faster: thresholds on the axis projection
        const int cc0 = (int)dot4(c[0], c[0]);
        const int k0 = (int)dot4(c[3], c[0]) - cc0;
        const int k1 = (int)dot4(c[3], c[1]) - (int)dot4(c[0], c[1]);
        const int k2 = (int)dot4(c[3], c[2]) - (int)dot4(c[0], c[2]);
        const int k3 = (int)dot4(c[3], c[3]) - (int)dot4(c[0], c[3]);
        const FasterCuts cut = faster_cuts(k0 + k1, k1 + k2, k2 + k3);
        const int x = d3 - d0;
        const bool b0 = x <= cut.f0, b1 = x <= cut.f1, b2 = x <= cut.f2;
        s = 3u - (uint32_t)b0 - (uint32_t)b1 - (uint32_t)b2;
        bu = b0 ? u0 : (b1 ? u1 : (b2 ? u2 : u3));
    }
    const int tot = (int)gsum((uint32_t)bu);
    const uint32_t sel = gor(s << (2 * glane()));
    return make_uint2(error_mode == QTX_ERROR_NONE ? 0u : (uint32_t)(cx.pp - tot), sel);
}

// One lane, all 16 texels from LDS, error only (a trial of cluster fit or search).
template <int INTERP>
__device__ __forceinline__ uint32_t trial_err(const Ctx &cx, int M, int error_mode, uint32_t ep) {
    uint32_t c[4];
    palette_pair<INTERP>(ep, (M & 0xF) == 3, c);
    const int cc0 = (int)dot4(c[0], c[0]), cc1 = (int)dot4(c[1], c[1]);
    const int cc2 = (int)dot4(c[2], c[2]), cc3 = (int)dot4(c[3], c[3]);
    const uint32_t n0 = neg_opaque((uint32_t)cc0), n1 = neg_opaque((uint32_t)cc1);
    const uint32_t n2 = neg_opaque((uint32_t)cc2), n3 = neg_opaque((uint32_t)cc3);
    int tot = 0;
    if (error_mode == QTX_ERROR_FULL) {
        if (M == QTX_COLOR_THREE) {
            for (int q = 0; q < 4; q++) {
                const uint4 t = tex4(q);
#pragma unroll
                for (int k = 0; k < 4; k++) {
                    const uint32_t p = pick(t, k);
                    tot += max(max(twice_plus(dot4(p, c[0]), n0), twice_plus(dot4(p, c[1]), n1)),
                               twice_plus(dot4(p, c[2]), n2));
                }
            }
        } else {
            for (int q = 0; q < 4; q++) {
                const uint4 t = tex4(q);
#pragma unroll
                for (int k = 0; k < 4; k++) {
                    const uint32_t p = pick(t, k);
                    const int u0 = twice_plus(dot4(p, c[0]), n0), u1 = twice_plus(dot4(p, c[1]), n1);
                    const int u2 = twice_plus(dot4(p, c[2]), n2), u3 = twice_plus(dot4(p, c[3]), n3);
                    tot += max(max(u0, u1), max(u2, u3));
                }
            }
        }
        return (uint32_t)(cx.pp - tot);
    }
    if (error_mode == QTX_ERROR_CHECK2) {
        const int a0 = (int)chn(c[3], 0) - (int)chn(c[0], 0);
        const int a1 = (int)chn(c[3], 1) - (int)chn(c[0], 1);
        const int a2 = (int)chn(c[3], 2) - (int)chn(c[0], 2);
        const uint32_t axis2 = (uint32_t)(a0 * a0 + a1 * a1 + a2 * a2);
        const Check2Cuts cut = check2_cuts(axis2, check2_f(axis2), (int)dot4(c[0], c[3]) - cc0);
        for (int q = 0; q < 4; q++) {
            const uint4 t = tex4(q);
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const uint32_t p = pick(t, k);
                const int d0 = (int)dot4(p, c[0]), d1 = (int)dot4(p, c[1]);
                const int d2 = (int)dot4(p, c[2]), d3 = (int)dot4(p, c[3]);
                const bool ge2 = d3 - d0 >= cut.t2, ge3 = d3 - d0 >= cut.t3;
                const int u0 = twice_plus((uint32_t)d0, n0), u1 = twice_plus((uint32_t)d1, n1);
                const int u2 = twice_plus((uint32_t)d2, n2), u3 = twice_plus((uint32_t)d3, n3);
                const int ua = ge2 ? (ge3 ? u2 : u1) : u0;
                const int ub = ge2 ? (ge3 ? u3 : u2) : u1;
                tot += max(ua, ub);
            }
        }
        return (uint32_t)(cx.pp - tot);
    }
    if (error_mode == QTX_ERROR_FASTER) {
        const int k0 = (int)dot4(c[3], c[0]) - cc0;
        const int k1 = (int)dot4(c[3], c[1]) - (int)dot4(c[0], c[1]);
        const int k2 = (int)dot4(c[3], c[2]) - (int)dot4(c[0], c[2]);
        const int k3 = cc3 - (int)dot4(c[0], c[3]);
        const FasterCuts cut = faster_cuts(k0 + k1, k1 + k2, k2 + k3);
        uint32_t pc = 0;
        for (int q = 0; q < 4; q++) {
            const uint4 t = tex4(q);
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const uint32_t p = pick(t, k);
                const int x = (int)dot4(p, c[3]) - (int)dot4(p, c[0]);
                const bool b0 = x <= cut.f0, b1 = x <= cut.f1, b2 = x <= cut.f2;
                const uint32_t cs = b0 ? c[0] : (b1 ? c[1] : (b2 ? c[2] : c[3]));
                tot += b0 ? cc0 : (b1 ? cc1 : (b2 ? cc2 : cc3));
                pc = dot4acc(p, cs, pc);
            }
        }
        return (uint32_t)(cx.pp - (int)(2u * pc) + tot);
    }
    return 0u;
}

// FindEndpointsSingleColor(result, color, is_3color): match-table endpoints for one
// colour (BC1Encoder.cpp:443-456); returns the pair, err gets the table error.
// The reference's 3-colour tables equal the 4-colour ones (SingleColorTable.h: the
// colour-count parameter is unused).
__device__ __forceinline__ uint32_t single_color(uint32_t color, uint32_t &err) {
    const uint32_t mr = L().match[0][chn(color, 0)];
    const uint32_t mg = L().match[1][chn(color, 1)];
    const uint32_t mb = L().match[0][chn(color, 2)];
    err = (mr >> 16) + (mg >> 16) + (mb >> 16);
    return ep_make((mr >> 8) & 0xFF, (mg >> 8) & 0xFF, (mb >> 8) & 0xFF) |
           (ep_make(mr & 0xFF, mg & 0xFF, mb & 0xFF) << 16);
}

// FindEndpointsSingleColor(result, pixels, color, is_3color) (BC1Encoder.cpp:458-471):
// match-table endpoints, error against colour 2 of the endpoints held on entry
// (alpha included), all selectors 1.  One lane, texels from LDS.
template <int INTERP> __device__ __forceinline__ Res single_color_px(const Ctx &cx, uint32_t ep_in, uint32_t color, bool three) {
    uint32_t c[4];
    palette_pair<INTERP>(ep_in, three, c);
    const uint32_t qc = c[1] | 0xFF000000u;  // colour 2 of the palette, alpha 255
    uint32_t pq = 0;
    for (int q = 0; q < 4; q++) {
        const uint4 t = tex4(q);
        pq += dot4(t.x, qc) + dot4(t.y, qc) + dot4(t.z, qc) + dot4(t.w, qc);
    }
    uint32_t e;
    Res r;
    r.ep = single_color(color, e);
    r.err = cx.ppa - 2u * pq + 16u * dot4(qc, qc);
    r.sel = 0x55555555u;
    r.mode = three ? QTX_COLOR_THREE : QTX_COLOR_FOUR;
    return r;
}

// RefineEndpointsLS(result, pixels, metrics) with 4-colour weights (BC1Encoder.cpp:781-825;
// RefineBlockLS always instantiates it <FourColor>, :860), texel-parallel.
// Returns 0xFFFFFFFF when the 2x2 system is singular.
__device__ __forceinline__ uint32_t ls_from_pixels(const Ctx &cx, uint32_t sel, const Metrics &m) {
    const uint32_t s = (sel >> (2 * glane())) & 3u;
    const uint32_t qa = gsum((chn(cx.px, 0) * s) | ((chn(cx.px, 1) * s) << 16));
    const uint32_t qb = gsum(chn(cx.px, 2) * s);
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
    return precise_round565_ls((a0 * q0) + (a1 * r0), (a0 * q1) + (a1 * r1), (a0 * q2) + (a1 * r2),
                               (a2 * q0) + (a3 * r0), (a2 * q1) + (a3 * r1), (a2 * q2) + (a3 * r2));
}

// RefineBlockLS<M> (BC1Encoder.cpp:852-875)
template <int INTERP>
__device__ __forceinline__ Res refine_block_ls(const Ctx &cx, const Metrics &m, int M, int error_mode, int passes, Res r) {
    for (int pass = 0; pass < passes; pass++) {
        Res t;
        const uint32_t ep = ls_from_pixels(cx, r.sel, m);
        if (ep == 0xFFFFFFFFu) {
            t = single_color_px<INTERP>(cx, r.ep, m.avg, M != QTX_COLOR_FOUR);
        } else {
            const uint2 e = eval_px<INTERP>(cx, M, error_mode, ep);
            t = Res{ep, e.y, e.x, M};
        }
        if (t.ep == r.ep) break;
        if (error_mode == QTX_ERROR_NONE || t.err < r.err) r = t;
        else break;
    }
    return r;
}

// FindEndpoints (BC1Encoder.cpp:473-677), texel-parallel.  Returns low | high << 16.
__device__ __forceinline__ uint32_t find_endpoints(const Ctx &cx, const Metrics &m, int endpoint_mode, bool ignore_black) {
    const uint32_t p = cx.px;
    const uint32_t pr = chn(p, 0), pg = chn(p, 1), pb = chn(p, 2);
    const uint32_t mn0 = chn(m.mn, 0), mn1 = chn(m.mn, 1), mn2 = chn(m.mn, 2);
    const uint32_t mx0 = chn(m.mx, 0), mx1 = chn(m.mx, 1), mx2 = chn(m.mx, 2);
    const int av0 = (int)chn(m.avg, 0), av1 = (int)chn(m.avg, 1), av2 = (int)chn(m.avg, 2);
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
        const uint32_t e0 = mx0 - mn0, e1 = mx1 - mn1, e2 = mx2 - mn2;
        const int chan0 = (e0 >= e1 && e0 >= e2) ? 0 : ((e1 >= e2 && e1 >= e0) ? 1 : 2);
        const int chan1 = (chan0 + 1) % 3, chan2 = (chan0 + 2) % 3;  // secondary axes, in the reference's order
        const uint32_t x = chn(p, chan0);
        const uint32_t sxx = gsum(x * x), sxy1 = gsum(x * chn(p, chan1)), sxy2 = gsum(x * chn(p, chan2));
        const int sums[3] = {m.s0, m.s1, m.s2};
        const uint32_t sum_x = (uint32_t)(chan0 == 0 ? sums[0] : (chan0 == 1 ? sums[1] : sums[2]));
        const uint32_t sum_y1 = (uint32_t)(chan1 == 0 ? sums[0] : (chan1 == 1 ? sums[1] : sums[2]));
        const uint32_t sum_y2 = (uint32_t)(chan2 == 0 ? sums[0] : (chan2 == 1 ? sums[1] : sums[2]));
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
        return precise_round565(lr, lg, lb) | (precise_round565(hr, hg, hb) << 16);
    }
    if (endpoint_mode == QTX_EP_BOUNDING_BOX || endpoint_mode == QTX_EP_BOUNDING_BOX_INT) {
        // covariance sign of r and g against b, with the reference's operator
        // precedence (:571-573, :596-598): sum of (b - avg.b) * r - avg.r
        const int bl = (int)pb - av2;
        const int icov_xz = (int)gsum((uint32_t)(bl * (int)pr - av0));
        const int icov_yz = (int)gsum((uint32_t)(bl * (int)pg - av1));
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
            return precise_round565(l0, l1, l2) | (precise_round565(h0, h1, h2) << 16);
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
    // PCA, :606-674.  Covariance = Identity + sum(diff diff^T) / 255: the products
    // and their sums are exact integers below 2^24, so any summation order gives
    // the reference's floats.
    const bool inc = !(ignore_black && is_black(p));
    const int d0i = inc ? (int)pr - av0 : 0, d1i = inc ? (int)pg - av1 : 0, d2i = inc ? (int)pb - av2 : 0;
    const float c00 = (1.0f + (float)(int)gsum((uint32_t)(d0i * d0i))) / 255.0f;
    const float c11 = (1.0f + (float)(int)gsum((uint32_t)(d1i * d1i))) / 255.0f;
    const float c22 = (1.0f + (float)(int)gsum((uint32_t)(d2i * d2i))) / 255.0f;
    const float c01 = (float)(int)gsum((uint32_t)(d0i * d1i)) / 255.0f;
    const float c02 = (float)(int)gsum((uint32_t)(d0i * d2i)) / 255.0f;
    const float c12 = (float)(int)gsum((uint32_t)(d1i * d2i)) / 255.0f;
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
    float dot = 0.0f;
    dot += (float)pr * ax0;
    dot += (float)pg * ax1;
    dot += (float)pb * ax2;
    dot += 0.0f * ax3;
    // first texel with the strictly smallest / largest projection (:662-669)
    uint32_t kmin = inc ? fkey(dot) : 0xFFFFFFFFu, imin = (uint32_t)glane();
    uint32_t kmax = inc ? ~fkey(dot) : 0xFFFFFFFFu, imax = (uint32_t)glane();
    gargmin(kmin, imin);
    gargmin(kmax, imax);
    return to565p(tex((int)imin)) | (to565p(tex((int)imax)) << 16);
}

// ---------------------------------------------------------------------------
// Cluster fit, RefineBlockCF<M> (BC1Encoder.cpp:877-928).
template <int INTERP>
__device__ __forceinline__ Res cluster_fit(const Ctx &cx, const Metrics &m, int M, int error_mode, uint32_t orderings,
                                           const Res orig) {
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
    // rank of this lane's texel by its projection on high - low (8-bit), ties by
    // texel index; the projection is an exact integer
    const uint32_t xh = expand565(ep_unpack(orig.ep >> 16)), xl = expand565(ep_unpack(orig.ep & 0xFFFFu));
    const uint32_t mykey = ((uint32_t)(0x1000000 + ((int)dot4(cx.px, xh) - (int)dot4(cx.px, xl))) << 4) | (uint32_t)lane;
    int rank = 0;
    for (int q = 0; q < 4; q++) {
        const uint4 t = tex4(q);
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const uint32_t v = pick(t, k);
            const uint32_t key = ((uint32_t)(0x1000000 + ((int)dot4(v, xh) - (int)dot4(v, xl))) << 4) | (uint32_t)(4 * q + k);
            rank += key < mykey;
        }
    }
    // lane k receives the texel of rank k; a scan leaves sums[k+1] in lane k (r | g << 16, and b)
    const int base = (int)(threadIdx.x & 63 & ~(G - 1));
    const uint32_t sa = gscan((uint32_t)__builtin_amdgcn_ds_permute((base + rank) * 4, (int)(chn(cx.px, 0) | (chn(cx.px, 1) << 16))));
    const uint32_t sb = gscan((uint32_t)__builtin_amdgcn_ds_permute((base + rank) * 4, (int)chn(cx.px, 2)));
    const uint32_t total_a = gshfl(sa, G - 1), total_b = gshfl(sb, G - 1);
    const float den = four ? 3.0f : 2.0f;
    const float s16_0 = (float)(total_a & 0xFFFFu), s16_1 = (float)(total_a >> 16), s16_2 = (float)total_b;

    const uint32_t q_total = S().exhaustive ? (four ? (uint32_t)kOrders4 : (uint32_t)kOrders3) : orderings;
    const uint16_t *best_row = four ? &cx.T->best4[start * kBestOrders4] : &cx.T->best3[start * kBestOrders3];
    uint32_t b_err = 0xFFFFFFFFu, b_q = 0xFFFFFFFFu, b_ep = 0;
    const uint32_t rounds = (q_total + G - 1) / G;
    for (uint32_t kk = 0; kk < rounds; kk++) {
        const uint32_t q = kk * G + (uint32_t)lane;
        const bool active = q < q_total;
        const uint32_t hash = active ? (S().exhaustive ? q : best_row[q]) : 0u;
        const uint32_t lv = four ? L().lv4[hash] : L().lv3[hash];
        const int l1 = (int)(lv & 0xFF), l2 = (int)((lv >> 8) & 0xFF), l3 = (int)((lv >> 16) & 0xFF);
        // prefix sums at the ordering's cumulative levels (every lane takes part)
        const uint32_t a1 = gshfl(sa, (l1 + 15) & 15), b1 = gshfl(sb, (l1 + 15) & 15);
        const uint32_t a2 = gshfl(sa, (l2 + 15) & 15), b2 = gshfl(sb, (l2 + 15) & 15);
        const uint32_t a3 = gshfl(sa, (l3 + 15) & 15), b3 = gshfl(sb, (l3 + 15) & 15);
        if (!active) continue;
        uint32_t ep, err;
        if (lv >> 24) {
            const Res t = single_color_px<INTERP>(cx, orig.ep, m.avg, !four);
            ep = t.ep;
            err = t.err;
        } else {
            // RefineEndpointsLS(result, sums, matrix, hash) (:827-850)
            const float4 fm = four ? L().f4[hash] : L().f3[hash];
            const uint32_t qa = (l1 ? a1 : 0u) + (l2 ? a2 : 0u) + ((four && l3) ? a3 : 0u);
            const uint32_t qb = (l1 ? b1 : 0u) + (l2 ? b2 : 0u) + ((four && l3) ? b3 : 0u);
            // q10 = sums[l1] + sums[l2] (+ sums[l3]), q00 = d * sums[16] - q10: exact integers
            const float q10_0 = (float)(qa & 0xFFFFu), q10_1 = (float)(qa >> 16), q10_2 = (float)qb;
            const float q00_0 = (s16_0 * den) - q10_0, q00_1 = (s16_1 * den) - q10_1, q00_2 = (s16_2 * den) - q10_2;
            ep = precise_round565_ls((fm.x * q00_0) + (fm.y * q10_0), (fm.x * q00_1) + (fm.y * q10_1),
                                     (fm.x * q00_2) + (fm.y * q10_2), (fm.z * q00_0) + (fm.w * q10_0),
                                     (fm.z * q00_1) + (fm.w * q10_1), (fm.z * q00_2) + (fm.w * q10_2));
            err = trial_err<INTERP>(cx, M, error_mode, ep);
        }
        if (err < b_err) {  // ties keep this lane's earliest ordering
            b_err = err;
            b_q = q;
            b_ep = ep;
        }
    }
    uint32_t e = b_err, q = b_q;
    gargmin(e, q);  // (error, ordering index), ties to the earliest ordering
    if (!(e < orig.err)) return orig;
    const uint32_t ep = gshfl(b_ep, (int)(q % G));
    const uint32_t hash = S().exhaustive ? q : best_row[q];
    const uint32_t lv = four ? L().lv4[hash] : L().lv3[hash];
    if (lv >> 24) return Res{ep, 0x55555555u, e, four ? QTX_COLOR_FOUR : QTX_COLOR_THREE};
    const uint2 w = eval_px<INTERP>(cx, M, error_mode, ep);  // the winner's selectors
    return Res{ep, w.y, e, M};
}

// ---------------------------------------------------------------------------
// EndpointSearch (BC1Encoder.cpp:930-996) with a speculative window of G rounds.
// Voxel i moves one endpoint by (dr, dg, db); voxel opp(i) is its reverse.
__device__ __forceinline__ int vox_d(uint32_t table, uint32_t v) { return (int)((table >> (2 * v)) & 3u) - 1; }
constexpr uint32_t vox_pack(const int (&d)[16]) {  // 2-bit fields, value + 1
    uint32_t r = 0;
    for (int i = 0; i < 16; i++) r |= (uint32_t)(d[i] + 1) << (2 * i);
    return r;
}
constexpr int kVoxR[16] = {1, 0, 0, -1, 0, 0, 1, 1, 0, -1, -1, 0, -1, 1, 0, 0};
constexpr int kVoxG[16] = {0, 1, 0, 0, -1, 0, 1, 0, 1, -1, 0, -1, 1, -1, -1, 1};
constexpr int kVoxB[16] = {0, 0, 1, 0, 0, -1, 0, 1, 1, 0, -1, -1, 0, 0, 1, -1};
constexpr int kVoxOppT[16] = {3, 4, 5, 0, 1, 2, 9, 10, 11, 6, 7, 8, 13, 12, 15, 14};
constexpr uint64_t opp_pack(const int (&o)[16]) {  // 4-bit fields
    uint64_t r = 0;
    for (int i = 0; i < 16; i++) r |= (uint64_t)o[i] << (4 * i);
    return r;
}
constexpr uint32_t kVoxRp = vox_pack(kVoxR), kVoxGp = vox_pack(kVoxG), kVoxBp = vox_pack(kVoxB);
constexpr uint64_t kVoxOpp = opp_pack(kVoxOppT);

template <int INTERP> __device__ __forceinline__ Res endpoint_search(const Ctx &cx, Res r) {
    const int lane = glane();
    const uint32_t R = S().search_rounds;
    uint32_t prev = 0, i0 = 0;
    int forbidden = -1;
    const bool three = r.mode == QTX_COLOR_THREE || r.mode == QTX_COLOR_THREE_BLACK;
    const int emode = three ? QTX_ERROR_FULL : S().error_mode;
    const int M = three ? r.mode : QTX_COLOR_FOUR;
    while (i0 < R) {
        // this lane's round: the lane-th round >= i0 that is not the forbidden direction
        const uint32_t f0 = forbidden >= 0 ? i0 + (((uint32_t)forbidden - (i0 & 31u)) & 31u) : 0xFFFFFFFFu;
        uint32_t i = i0 + (uint32_t)lane;
        if (i >= f0) i += 1;
        uint32_t ip = i0 + (uint32_t)lane - 1u;  // round of the previous lane
        if (lane > 0 && ip >= f0) ip += 1;
        // evaluated by the sequential loop unless an earlier round ended it
        const bool valid = i < R && (lane == 0 || ip <= prev + 32u);
        uint32_t tep = r.ep, terr = 0xFFFFFFFFu;
        if (valid) {
            const uint32_t v = i & 15u;
            const uint32_t ep = (i & 16u) ? (r.ep & 0xFFFFu) : (r.ep >> 16);
            const uint32_t er = (uint32_t)clampi((int)ep_r(ep) + vox_d(kVoxRp, v), 0, 31);
            const uint32_t eg = (uint32_t)clampi((int)ep_g(ep) + vox_d(kVoxGp, v), 0, 63);
            const uint32_t eb = (uint32_t)clampi((int)ep_b(ep) + vox_d(kVoxBp, v), 0, 31);
            const uint32_t ne = ep_make(er, eg, eb);
            tep = (i & 16u) ? ((r.ep & 0xFFFF0000u) | ne) : ((r.ep & 0xFFFFu) | (ne << 16));
            terr = trial_err<INTERP>(cx, M, emode, tep);
        }
        const uint32_t imp = gballot(valid && terr < r.err);
        if (imp) {
            const int j = __builtin_ctz(imp);
            const uint32_t ij = gshfl(i, j);
            const uint32_t nep = gshfl(tep, j);
            const uint2 w = eval_px<INTERP>(cx, M, emode, nep);
            r = Res{nep, w.y, w.x, M};
            forbidden = (int)((kVoxOpp >> (4 * (ij & 15u))) & 15u) | (int)(ij & 16u);
            prev = ij;
            i0 = ij + 1;
            continue;
        }
        // no improvement in the window: stop if it ran past the last round or the
        // loop's exit test, else continue after its last round
        const uint32_t ilast = gshfl(i, G - 1);
        if (gballot(valid) != 0xFFFFu || ilast > prev + 32u) break;
        i0 = ilast + 1;
    }
    return r;
}

// ---------------------------------------------------------------------------
// Block writers (BC1Encoder.cpp:345-441), 8 bytes as two little-endian words.

__device__ __forceinline__ uint2 write_solid(uint32_t color) {
    uint32_t mask = 0xAA, min16, max16;
    if ((chn(color, 0) | chn(color, 1) | chn(color, 2)) == 0) {
        min16 = 0; max16 = 1; mask = 0x55;
    } else {
        uint32_t e;
        const uint32_t ep = single_color(color, e);
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
    uint32_t ep1 = r.ep & 0xFFFFu, ep0 = r.ep >> 16;
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
// BC1Encoder::EncodeBlock (BC1Encoder.cpp:261-342) for one group.
template <int INTERP> __device__ __forceinline__ uint2 encode_bc1(const Ctx &cx) {
    const uint32_t first = tex(0);
    if (gballot(cx.px != first) == 0) return write_solid(first);  // alpha included (ColorBlock.h:80-86)

    const Metrics m = metrics(cx, false);
    const BC1Settings &s = S();
    const int error_mode = s.error_mode;  // orderings >= 1 always makes block errors necessary

    Res orig{0u, 0u, 0xFFFFFFFFu, 0}, result{0u, 0u, 0xFFFFFFFFu, 0};
    for (int round = 0; round < s.ep_passes; round++) {
        const int ep_mode = round == 1 ? QTX_EP_BOUNDING_BOX : s.endpoint_mode;
        const uint32_t ep = find_endpoints(cx, m, ep_mode, false);
        const uint2 e = eval_px<INTERP>(cx, QTX_COLOR_FOUR, error_mode, ep);
        const Res t = refine_block_ls<INTERP>(cx, m, QTX_COLOR_FOUR, error_mode, s.ls_passes, Res{ep, e.y, e.x, QTX_COLOR_FOUR});
        if (t.err < result.err) {
            result = t;
            orig = Res{ep, 0u, 0xFFFFFFFFu, 0};
        }
    }
    if (result.err > 0) {
        for (int it = 0; it < s.cf_passes; it++)
            result = cluster_fit<INTERP>(cx, m, QTX_COLOR_FOUR, s.error_mode, s.orderings4, result);
    }
    if (result.err > 0 && (s.color_mode & QTX_COLOR_THREE)) {
        const uint2 e = eval_px<INTERP>(cx, QTX_COLOR_THREE, QTX_ERROR_FULL, orig.ep);
        Res t = refine_block_ls<INTERP>(cx, m, QTX_COLOR_THREE, QTX_ERROR_FULL, s.ls_passes, Res{orig.ep, e.y, e.x, QTX_COLOR_THREE});
        if (t.err > 0) {
            for (int it = 0; it < s.cf_passes; it++)
                t = cluster_fit<INTERP>(cx, m, QTX_COLOR_THREE, QTX_ERROR_FULL, s.orderings3, t);
        }
        if (t.err < result.err) result = t;
    }
    if (result.err > 0 && s.color_mode == QTX_COLOR_THREE_BLACK && m.has_black &&
        !((chn(m.mx, 0) | chn(m.mx, 1) | chn(m.mx, 2)) < 4u)) {
        const Metrics mnb = metrics(cx, true);
        const uint32_t ep = find_endpoints(cx, mnb, QTX_EP_PCA, true);
        const uint2 e = eval_px<INTERP>(cx, QTX_COLOR_THREE_BLACK, QTX_ERROR_FULL, ep);
        const Res t = refine_block_ls<INTERP>(cx, mnb, QTX_COLOR_THREE_BLACK, QTX_ERROR_FULL, s.ls_passes,
                                              Res{ep, e.y, e.x, QTX_COLOR_THREE_BLACK});
        if (t.err < result.err) result = t;
    }
    if (result.err > 0 && s.search_rounds > 0) result = endpoint_search<INTERP>(cx, result);
    return write_block(result);
}

// Copy the launch's tables into LDS and s into every wave's settings slot (every
// thread of the workgroup takes part).
template <int INTERP> __device__ __forceinline__ void load_lds_tables(const DeviceTables *T, const BC1Settings &s) {
    Lds &l = L();
    for (int i = (int)threadIdx.x; i < kOrders4; i += (int)blockDim.x) {
        l.f4[i] = *reinterpret_cast<const float4 *>(T->factors4[i]);
        l.lv4[i] = T->levels4[i];
    }
    for (int i = (int)threadIdx.x; i < kOrders3; i += (int)blockDim.x) {
        l.f3[i] = *reinterpret_cast<const float4 *>(T->factors3[i]);
        l.lv3[i] = T->levels3[i];
    }
    for (int i = (int)threadIdx.x; i < 512; i += (int)blockDim.x) l.match[i >> 8][i & 255] = T->match[INTERP][i >> 8][i & 255];
    if (threadIdx.x < 32) l.mid5[threadIdx.x] = T->mid5[threadIdx.x];
    if (threadIdx.x < 64) l.mid6[threadIdx.x] = T->mid6[threadIdx.x];
    if (threadIdx.x < kWaves) l.s[threadIdx.x] = s;
}

}  // namespace dev
}  // namespace qtx
