// resample.hip -- gfx950 kernels of the mip chain (Pillow's BILINEAR resize per
// channel, from the base image) and their host plan.  The algorithm is stated in
// resample.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "resample.h"

namespace qtx {
namespace dev {

constexpr int kBits = 22;  // fraction bits of the fixed-point weights

// Window and weights of output index i of an axis of in_size -> out_size
// (libImaging/Resample.c, precompute_coeffs + normalize_coeffs_8bpc, bilinear
// filter): the same double operations in the same order as Pillow.  The sum of
// the weights is sequential, in tap order, as Pillow's (double addition is not
// associative); everything else is per tap, so a wave shares a long window:
// `lanes` lanes with ids `lane` (one lane for a short window, 64 for a long one).
struct Window {
    double center, ss;
    int xmin, n;
    __device__ double weight(int x) const {
        double t = (x + xmin - center + 0.5) * ss;
        if (t < 0.0) t = -t;
        return t < 1.0 ? 1.0 - t : 0.0;
    }
};

__device__ Window axis_window(int in_size, int out_size, int i) {
    const double scale = (double)in_size / (double)out_size;
    const double filterscale = scale < 1.0 ? 1.0 : scale;
    const double support = 1.0 * filterscale;
    Window wd;
    wd.center = 0.0 + (i + 0.5) * scale;
    wd.ss = 1.0 / filterscale;
    int xmin = (int)(wd.center - support + 0.5);
    if (xmin < 0) xmin = 0;
    int xmax = (int)(wd.center + support + 0.5);
    if (xmax > in_size) xmax = in_size;
    wd.xmin = xmin;
    wd.n = xmax - xmin;
    return wd;
}

// The double of lane j (j uniform across the wave).
__device__ __forceinline__ double lane_double(double v, int j) {
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), j);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), j);
    return __hiloint2double(hi, lo);
}

// Weights of output index i, tap x stored at k[x * stride] (tables are tap-major,
// so that the lanes of a pass, on consecutive outputs, read consecutive words).
// lanes == 1: one lane does everything.  lanes == 64: the wave computes 64 taps
// at a time and every lane adds them to the sum in tap order (each value read
// from its lane), so the sum is Pillow's sequential one.
__device__ void axis_coeffs(int in_size, int out_size, int i, int lane, int lanes, bool exact,
                            int32_t *__restrict__ k, int stride, int32_t *__restrict__ win) {
    const Window wd = axis_window(in_size, out_size, i);
    double ww = 0.0;
    if (lanes == 1) {
        for (int x = 0; x < wd.n; x++) ww += wd.weight(x);
    } else if (exact) {
        // a power-of-two scale: every weight is a dyadic fraction with at most
        // log2(scale) + 1 fraction bits and every partial sum is exact, so the
        // sum does not depend on the order of the additions
        for (int x = lane; x < wd.n; x += 64) ww += wd.weight(x);
#pragma unroll
        for (int d = 32; d >= 1; d >>= 1) ww += __shfl_xor(ww, d, 64);
    } else {
        for (int b = 0; b < wd.n; b += 64) {
            const double wx = b + lane < wd.n ? wd.weight(b + lane) : 0.0;
            const int cnt = min(64, wd.n - b);
            for (int j = 0; j < cnt; j++) ww += lane_double(wx, j);
        }
    }
    for (int x = lane; x < wd.n; x += lanes) {
        double v = wd.weight(x);
        if (ww != 0.0) v /= ww;
        k[(size_t)x * stride] = v < 0.0 ? (int32_t)(-0.5 + v * (1 << kBits)) : (int32_t)(0.5 + v * (1 << kBits));
    }
    if (lane == 0) {
        win[0] = wd.xmin;
        win[1] = wd.n;
    }
}

// Level of a launch-wide index: the last level whose first index is <= u.
__device__ __forceinline__ int level_of(const uint32_t *first, int levels, uint32_t u) {
    int m = 0;
    for (int l = 1; l < levels; l++)
        if (first[l] <= u) m = l;
    return m;
}

// Coefficients of every level and axis, in waves: an axis whose window is at most
// kResampleWideTaps taps long takes 64 output indices per wave (one per lane), a
// longer one one index per wave.  coef0 counts waves: a level's x waves, then its
// y waves.
__device__ __forceinline__ uint32_t axis_waves(int out_size, int ks) {
    return ks > kResampleWideTaps ? (uint32_t)out_size : (uint32_t)((out_size + 63) / 64);
}

__global__ __launch_bounds__(256) void k_mip_coeffs(MipPlan p, int32_t *__restrict__ ws) {
    const uint32_t wave = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = (int)(threadIdx.x & 63);
    if (wave >= p.coef0[p.levels]) return;  // wave-uniform
    const int m = level_of(p.coef0, p.levels, wave);
    uint32_t local = wave - p.coef0[m];
    const uint32_t xw = axis_waves(p.lw[m], p.ksx[m]);
    const bool yaxis = local >= xw;
    if (yaxis) local -= xw;
    const int in_size = yaxis ? p.h : p.w, out_size = yaxis ? p.lh[m] : p.lw[m];
    const int ks = yaxis ? p.ksy[m] : p.ksx[m];
    int32_t *kt = ws + (yaxis ? p.cy[m] : p.cx[m]);
    int32_t *wt = ws + (yaxis ? p.wy[m] : p.wx[m]);
    if (ks > kResampleWideTaps) {
        const int i = (int)local;
        const bool exact = ((yaxis ? p.exact_y : p.exact_x) >> m) & 1u;
        axis_coeffs(in_size, out_size, i, lane, 64, exact, kt + i, out_size, wt + 2 * i);
    } else {
        const int i = (int)local * 64 + lane;
        if (i < out_size) axis_coeffs(in_size, out_size, i, 0, 1, false, kt + i, out_size, wt + 2 * i);
    }
}

// Weights of the triangle filter are never negative and at most 2^22 and texels
// at most 255, so every product fits the 24-bit multiplier (v_mad_u32_u24, full
// rate; a 32-bit multiply is a quarter-rate instruction) and every sum, at most
// 2^21 + 255 * (2^22 + 64), fits 31 bits.
__device__ __forceinline__ void accumulate(uint32_t px, int32_t k, int32_t (&ss)[4]) {
#pragma unroll
    for (int c = 0; c < 4; c++) ss[c] += (int32_t)__umul24((px >> (8 * c)) & 0xFFu, (uint32_t)k);
}

// The four clamped channels packed as RGBA8.  The optimisation barrier keeps the
// compiler from fusing shift, clamp and packing of two channels into
// v_ashr_pk_u8_i32: this toolchain emits that instruction as if it cleared the
// upper half of its destination, which on gfx950 keeps its previous bits, and
// those then leak into the third and fourth bytes.
__device__ __forceinline__ uint32_t pack_clip(const int32_t (&ss)[4]) {
    uint32_t v = 0;
#pragma unroll
    for (int c = 0; c < 4; c++) {
        const int32_t s = ss[c] >> kBits;
        uint32_t q = (uint32_t)(s < 0 ? 0 : s > 255 ? 255 : s);
        asm volatile("" : "+v"(q));
        v |= q << (8 * c);
    }
    return v;
}

// Sum over aligned groups of G lanes (G = 1, 2, 4, ..., 64, uniform) with DPP:
// quad swaps and row mirrors leave each 2/4/8/16-lane group's sum in all its
// lanes; row broadcasts then add rows 0+1 into row 1 and 2+3 into row 3 (G = 32)
// and rows 0-1 into rows 2-3 (G = 64).  The last lane of every group holds the sum.
template <int CTRL, int ROWS = 0xF> __device__ __forceinline__ int32_t dpp_add(int32_t v) {
    return v + __builtin_amdgcn_update_dpp(0, v, CTRL, ROWS, 0xF, false);
}
__device__ __forceinline__ int32_t group_sum(int32_t v, int G) {
    if (G >= 2) v = dpp_add<0xB1>(v);         // quad_perm [1,0,3,2]
    if (G >= 4) v = dpp_add<0x4E>(v);         // quad_perm [2,3,0,1]
    if (G >= 8) v = dpp_add<0x141>(v);        // row_half_mirror
    if (G >= 16) v = dpp_add<0x140>(v);       // row_mirror
    if (G >= 32) v = dpp_add<0x142, 0xA>(v);  // row_bcast:15 into rows 1 and 3
    if (G >= 64) v = dpp_add<0x143, 0xC>(v);  // row_bcast:31 into rows 2 and 3
    return v;
}

// Sum of the four channel sums over the wave (integer: order-free).
__device__ __forceinline__ void wave_sum(int32_t (&ss)[4]) {
#pragma unroll
    for (int c = 0; c < 4; c++)
#pragma unroll
        for (int d = 32; d >= 1; d >>= 1) ss[c] += __shfl_xor(ss[c], d, 64);
}

// One pass of every level that needs it.  VERTICAL = false: base (w x h) ->
// lw x h; true: (lw x h, the horizontal result or the base) -> lw x lh.  A level
// whose window holds at most kResampleWideTaps taps takes one output pixel per
// lane (256 per workgroup), a longer one one output pixel per wave (4 per
// workgroup, lanes striding over the taps).
template <bool VERTICAL>
__global__ __launch_bounds__(256) void k_mip_pass(MipPlan p, const uint8_t *__restrict__ base, uint8_t *__restrict__ ws) {
    const uint32_t *first = VERTICAL ? p.v0 : p.h0;
    const uint32_t u = blockIdx.x;
    const int m = level_of(first, p.levels, u);
    const int lw = p.lw[m], lh = p.lh[m];
    const int ks = VERTICAL ? p.ksy[m] : p.ksx[m];
    const int32_t *kt = reinterpret_cast<const int32_t *>(ws) + (VERTICAL ? p.cy[m] : p.cx[m]);
    const int32_t *win = reinterpret_cast<const int32_t *>(ws) + (VERTICAL ? p.wy[m] : p.wx[m]);
    const bool both = lw != p.w && lh != p.h;
    const uint32_t *src;
    uint32_t *dst;
    int sw;  // source row length in pixels
    if (VERTICAL) {
        src = reinterpret_cast<const uint32_t *>(lw != p.w ? ws + p.tmp[m] : base);
        sw = lw;
        dst = reinterpret_cast<uint32_t *>(p.dst[m]);
    } else {
        src = reinterpret_cast<const uint32_t *>(base);
        sw = p.w;
        dst = reinterpret_cast<uint32_t *>(both ? ws + p.tmp[m] : p.dst[m]);
    }
    const int out_h = VERTICAL ? lh : p.h;
    const uint64_t total = (uint64_t)lw * (uint64_t)out_h;
    const bool wide = ks > kResampleWideTaps;
    const uint64_t o = wide ? (uint64_t)(u - first[m]) * 4u + (threadIdx.x >> 6)
                            : (uint64_t)(u - first[m]) * 256u + threadIdx.x;
    if (o >= total) return;  // wave-uniform when wide
    const int x = (int)(o % (uint64_t)lw), y = (int)(o / (uint64_t)lw);
    const int i = VERTICAL ? y : x;  // index along the pass's axis
    const int lo = win[2 * i], n = win[2 * i + 1];
    const int32_t *k = kt + i;
    const size_t kstride = VERTICAL ? (size_t)lh : (size_t)lw;  // tap-major tables
    int32_t ss[4] = {0, 0, 0, 0};
    const int t0 = wide ? (int)(threadIdx.x & 63) : 0, dt = wide ? 64 : 1;
    for (int t = t0; t < n; t += dt) {
        const uint32_t px = VERTICAL ? src[(size_t)(lo + t) * sw + x] : src[(size_t)y * sw + lo + t];
        accumulate(px, k[t * kstride], ss);
    }
    if (wide) wave_sum(ss);
    if (!wide || (threadIdx.x & 63) == 0) {
#pragma unroll
        for (int c = 0; c < 4; c++) ss[c] += 1 << (kBits - 1);
        dst[(size_t)y * lw + x] = pack_clip(ss);
    }
}

// The horizontal pass of every level for whole base rows: a workgroup copies a
// row into LDS once and computes that row of every level from it.  Each output
// of level m is summed by a group of G = gx[m] lanes (1; 16 for windows of 32
// taps and more; 64 for windows of 512 and more), lane l of the group taking taps
// l, l + G, ...; the group adds its partial sums with DPP lane moves (integer
// sums: any order gives the same total).
// Rows of at most kRowMaxWidth texels.
//
// Box sums.  A shrink by r = 2^a (16 <= r <= 1024) gives an interior output o the
// window [(2o - 1) h, (2o + 3) h), h = r / 2, with the weights (2t + 1) << sh
// rising over its first 2h taps and (4r - 2t - 1) << sh falling over the last
// 2h (sh = 21 - 2a; see k_mip_hrows).  Over the four boxes of h texels it spans,
// q = 2o - 1 .. 2o + 2, with S = the box's texel sum and M = the sum of its texels
// times their offset in the box, that weighted sum is exactly
//   2^sh (2 M_q + S_q + 2 M_q+1 + (2h + 1) S_q+1 + (4h - 1) S_q+2 - 2 M_q+2
//         + (2h - 1) S_q+3 - 2 M_q+3),
// so such an output costs a handful of operations whatever its window, once the
// row's box sums are known; those of width 2h come from those of width h (S as
// the sum of the two halves, M as M_lo + M_hi + h S_hi), level after level.  All
// terms are non-negative integers below 2^31: the weighted sum is at most
// 255 x 2^22.  The sums live in LDS after the row, per box S of the four channels
// then M of the four (boxes_x marks the levels that use them; rows of at most
// kBoxMaxWidth texels).
extern __shared__ __align__(16) uint32_t mip_lds[];

// Box sums of width 8 straight from the row: box b of nb at out[c * nb + b] (S) and
// out[(4 + c) * nb + b] (M), c the channel.
__device__ __forceinline__ void boxes_from_row(const uint32_t *__restrict__ row, int nb, int32_t *__restrict__ out) {
    for (int b = (int)threadIdx.x; b < nb; b += 256) {
        int32_t ss[4] = {0, 0, 0, 0}, mm[4] = {0, 0, 0, 0};
#pragma unroll
        for (int i = 0; i < 8; i++) {
            const uint32_t px = row[8 * b + i];
#pragma unroll
            for (int c = 0; c < 4; c++) {
                const int32_t v = (int32_t)((px >> (8 * c)) & 0xFFu);
                ss[c] += v;
                mm[c] += i * v;
            }
        }
#pragma unroll
        for (int c = 0; c < 4; c++) {
            out[c * nb + b] = ss[c];
            out[(4 + c) * nb + b] = mm[c];
        }
    }
}

// Box sums of width 2h (nb boxes) from those of width h (2 nb boxes).
__device__ __forceinline__ void boxes_double(const int32_t *__restrict__ in, int h, int nb, int32_t *__restrict__ out) {
    for (int b = (int)threadIdx.x; b < nb; b += 256) {
#pragma unroll
        for (int c = 0; c < 4; c++) {
            const int32_t slo = in[c * 2 * nb + 2 * b], shi = in[c * 2 * nb + 2 * b + 1];
            const int32_t mlo = in[(4 + c) * 2 * nb + 2 * b], mhi = in[(4 + c) * 2 * nb + 2 * b + 1];
            out[c * nb + b] = slo + shi;
            out[(4 + c) * nb + b] = mlo + mhi + h * shi;
        }
    }
}

// Interior output o of a shrink by r = 2h = 2^a from the box sums of width h.
__device__ __forceinline__ uint32_t box_output(const int32_t *__restrict__ bx, int nb, int h, int sh, int o) {
    const int q = 2 * o - 1;
    int32_t ss[4];
#pragma unroll
    for (int c = 0; c < 4; c++) {
        const int32_t *S = bx + c * nb + q, *M = bx + (4 + c) * nb + q;
        const int32_t t = 2 * M[0] + S[0] + 2 * M[1] + (2 * h + 1) * S[1] + (4 * h - 1) * S[2] - 2 * M[2] +
                          (2 * h - 1) * S[3] - 2 * M[3];
        ss[c] = (t << sh) + (1 << (kBits - 1));
    }
    return pack_clip(ss);
}

__global__ __launch_bounds__(256) void k_mip_hrows(MipPlan p, const uint8_t *__restrict__ base,
                                                   uint8_t *__restrict__ ws) {
    uint32_t *row = mip_lds;
    // box sums, two buffers used in turn: width 8 (p.w words), then 16 (p.w / 2)
    int32_t *const boxbuf[2] = {reinterpret_cast<int32_t *>(mip_lds + p.w),
                                reinterpret_cast<int32_t *>(mip_lds + 2 * p.w)};
    const int32_t *wsi = reinterpret_cast<const int32_t *>(ws);
    const int tid = (int)threadIdx.x;
    // the workgroup's next row is read into registers while it works on the
    // current one (kRowPrefetch texels per thread; wider rows read the rest then)
    constexpr int kRowPrefetch = kBoxMaxWidth / 256;
    uint32_t ahead[kRowPrefetch];
    auto fetch = [&](int y) {
        const uint32_t *src = reinterpret_cast<const uint32_t *>(base) + (size_t)y * p.w;
#pragma unroll
        for (int k = 0; k < kRowPrefetch; k++) {
            const int x = tid + 256 * k;
            ahead[k] = x < p.w ? src[x] : 0u;
        }
    };
    if ((int)blockIdx.x < p.h) fetch((int)blockIdx.x);
    for (int y = (int)blockIdx.x; y < p.h; y += (int)gridDim.x) {
        const uint32_t *src = reinterpret_cast<const uint32_t *>(base) + (size_t)y * p.w;
#pragma unroll
        for (int k = 0; k < kRowPrefetch; k++) {
            const int x = tid + 256 * k;
            if (x < p.w) row[x] = ahead[k];
        }
        for (int x = tid + 256 * kRowPrefetch; x < p.w; x += 256) row[x] = src[x];
        __syncthreads();
        if (y + (int)gridDim.x < p.h) fetch(y + (int)gridDim.x);
        int box_h = 0, box_cur = 1;  // width of the box sums built so far (0: none) and their buffer
        for (int m = 1; m < p.levels; m++) {
            const int lw = p.lw[m];
            if (lw == p.w) continue;
            const int G = p.gx[m], gshift = __builtin_ctz((unsigned)G);
            const int32_t *kt = wsi + p.cx[m];
            const int32_t *win = wsi + p.wx[m];
            // A shrink by r = 2^a (a <= 10): an interior window holds 2r taps whose
            // weights are (t + 1/2) / r^2 rising and mirrored falling -- Pillow's
            // (2t + 1) << (21 - 2a) in 22-bit fixed point, exactly -- so those
            // windows take their weights from t instead of the table.
            const int r = p.w / lw;
            const bool closed = ((p.exact_x >> m) & 1u) && r <= 1024;
            const int sh = closed ? 21 - 2 * __builtin_ctz((unsigned)r) : 0;
            uint32_t *dst = reinterpret_cast<uint32_t *>(p.lh[m] != p.h ? ws + p.tmp[m] : p.dst[m]) + (size_t)y * lw;
            // interior outputs 1 .. lw - 2 from box sums of width r / 2 (uniform)
            const bool boxes = ((p.boxes_x >> m) & 1u) && (r / 2 == 8 || r / 2 == 2 * box_h);
            if (boxes) {
                const int h = r / 2, nb = p.w / h;
                int32_t *next = boxbuf[box_cur ^ 1];
                if (h == 8) boxes_from_row(row, nb, next);
                else boxes_double(boxbuf[box_cur], h / 2, nb, next);
                __syncthreads();
                box_cur ^= 1;
                box_h = h;
                for (int o = 1 + tid; o < lw - 1; o += 256) dst[o] = box_output(next, nb, h, sh, o);
            }
            // the rest: every output, or with box sums only the two edge outputs
            const int nout = boxes ? 2 : lw;
            const int lanes = nout * G;
            for (int b = 0; b < lanes; b += 256) {  // uniform across the workgroup
                const int it = b + tid;
                const int k = it >> gshift, l = it & (G - 1);
                const int o = boxes ? (k == 0 ? 0 : lw - 1) : k;
                const bool act = k < nout;  // uniform within a group
                int32_t ss[4] = {0, 0, 0, 0};
                if (act) {
                    const int lo = win[2 * o], n = win[2 * o + 1];
                    if (closed && n == 2 * r) {
#pragma unroll 8
                        for (int t = l; t < n; t += G)
                            accumulate(row[lo + t], (t < r ? 2 * t + 1 : 4 * r - 2 * t - 1) << sh, ss);
                    } else {
                        // unrolled so that several coefficient loads are in flight
#pragma unroll 8
                        for (int t = l; t < n; t += G) accumulate(row[lo + t], kt[(size_t)t * lw + o], ss);
                    }
                }
#pragma unroll
                for (int c = 0; c < 4; c++) ss[c] = group_sum(ss[c], G);  // every lane takes part
                if (act && l == G - 1) {
#pragma unroll
                    for (int c = 0; c < 4; c++) ss[c] += 1 << (kBits - 1);
                    dst[o] = pack_clip(ss);
                }
            }
        }
        __syncthreads();  // the next row overwrites the LDS
    }
}

// ---------------------------------------------------------------------------
// Per-texel helpers of the DDS encoder's callers (quicktex/cli/encode.py).

// *flag = 1 when a texel's alpha is below 255 (the alpha histogram test of
// `encode auto`, cli/encode.py:129-133: a count in any bin but 255).  16 bytes
// per lane; the tail texels one by one.  Writers only ever store 1.
__global__ __launch_bounds__(256) void k_has_alpha(const uint8_t *__restrict__ rgba, uint64_t texels,
                                                   uint32_t *__restrict__ flag) {
    const uint64_t quads = texels / 4;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    bool below = false;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < quads; i += stride) {
        const uint4 q = reinterpret_cast<const uint4 *>(rgba)[i];
        below |= (q.x & q.y & q.z & q.w) < 0xFF000000u;
    }
    const uint64_t t = quads * 4 + (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < texels) below |= reinterpret_cast<const uint32_t *>(rgba)[t] < 0xFF000000u;
    if (__ballot(below) != 0 && (threadIdx.x & 63) == 0) *flag = 1u;
}

// The normal-map swizzle of `encode bc3nm` (cli/encode.py:52-55): each texel
// (R, G, B, A) becomes (255, G, G, R).
__global__ __launch_bounds__(256) void k_swizzle_bc3nm(uint8_t *__restrict__ rgba, uint64_t texels) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    uint32_t *px = reinterpret_cast<uint32_t *>(rgba);
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < texels; i += stride) {
        const uint32_t v = px[i];
        const uint32_t r = v & 0xFFu, g = (v >> 8) & 0xFFu;
        px[i] = 0xFFu | (g << 8) | (g << 16) | (r << 24);
    }
}

}  // namespace dev

int mip_level_count(int32_t w, int32_t h, int32_t requested) {
    int n = 1;
    while ((w > 1 || h > 1) && (requested <= 0 || n < requested)) {
        w = std::max(w / 2, 1);
        h = std::max(h / 2, 1);
        n++;
    }
    return n;
}

namespace {

// Coefficient slots per output index (libImaging/Resample.c: ceil(support) * 2 + 1).
int axis_ksize(int in_size, int out_size) {
    const double scale = (double)in_size / (double)out_size;
    const double support = 1.0 * (scale < 1.0 ? 1.0 : scale);
    return (int)std::ceil(support) * 2 + 1;
}

uint64_t align256(uint64_t v) { return (v + 255) & ~(uint64_t)255; }

// in / out is a power of two of at most 2^25 (the weights then stay exact dyadic
// fractions, see axis_coeffs)
bool power_of_two_scale(int in_size, int out_size) {
    if (in_size % out_size) return false;
    const int r = in_size / out_size;
    return (r & (r - 1)) == 0 && r <= (1 << 25);
}

// Waves of the coefficient launch for one axis (k_mip_coeffs).
uint32_t coef_waves(int out_size, int ks) {
    return ks > kResampleWideTaps ? (uint32_t)out_size : (uint32_t)((out_size + 63) / 64);
}

uint32_t units(uint64_t outputs, int ks) {
    const uint64_t per = ks > kResampleWideTaps ? 4 : 256;
    return (uint32_t)((outputs + per - 1) / per);
}

}  // namespace

size_t mip_plan(int32_t w, int32_t h, int32_t levels, MipPlan &p) {
    p = MipPlan{};
    p.w = w;
    p.h = h;
    p.levels = levels;
    uint64_t words = 0;  // int32 words of coefficient and window tables
    for (int m = 0; m < levels; m++) {
        p.lw[m] = m ? std::max(p.lw[m - 1] / 2, 1) : w;
        p.lh[m] = m ? std::max(p.lh[m - 1] / 2, 1) : h;
        p.ksx[m] = axis_ksize(w, p.lw[m]);
        p.ksy[m] = axis_ksize(h, p.lh[m]);
        if (power_of_two_scale(w, p.lw[m])) p.exact_x |= 1u << m;
        // one lane per output for short windows (consecutive outputs read LDS words
        // `scale` apart: at most 8-way bank conflicts), a 16-lane group for long
        // ones (a group reads 16 consecutive words), a whole wave for windows of 512
        // taps and more (the few outputs of the last levels, whose taps would
        // otherwise queue up on a handful of lanes)
        const double scale_x = (double)w / (double)p.lw[m];
        p.gx[m] = scale_x >= 256.0 ? 64 : scale_x >= 16.0 ? 16 : 1;
        if (power_of_two_scale(h, p.lh[m])) p.exact_y |= 1u << m;
        // box sums for the interior outputs of power-of-two shrinks by 16 .. 1024
        // (below 16 the taps cost no more than the sums), on rows that fit
        const int r = w / p.lw[m];
        if (m > 0 && ((p.exact_x >> m) & 1u) && r >= 16 && r <= 1024 && p.lw[m] >= 3 && w <= kBoxMaxWidth)
            p.boxes_x |= 1u << m;
    }
    uint32_t coef = 0, hu = 0, vu = 0;
    for (int m = 0; m < levels; m++) {
        const bool need = m > 0;  // level 0 is the base itself
        p.coef0[m] = coef;
        p.h0[m] = hu;
        p.v0[m] = vu;
        if (!need) continue;
        p.cx[m] = words;
        words += (uint64_t)p.lw[m] * p.ksx[m];
        p.cy[m] = words;
        words += (uint64_t)p.lh[m] * p.ksy[m];
        p.wx[m] = words;
        words += 2ull * p.lw[m];
        p.wy[m] = words;
        words += 2ull * p.lh[m];
        coef += coef_waves(p.lw[m], p.ksx[m]) + coef_waves(p.lh[m], p.ksy[m]);
        if (p.lw[m] != w) hu += units((uint64_t)p.lw[m] * h, p.ksx[m]);
        if (p.lh[m] != h) vu += units((uint64_t)p.lw[m] * p.lh[m], p.ksy[m]);
    }
    p.coef0[levels] = coef;
    p.h0[levels] = hu;
    p.v0[levels] = vu;
    uint64_t bytes = align256(words * 4);
    for (int m = 1; m < levels; m++) {
        if (p.lw[m] != w && p.lh[m] != h) {
            p.tmp[m] = bytes;
            bytes = align256(bytes + (uint64_t)p.lw[m] * h * 4);
        }
    }
    return (size_t)bytes;
}

hipError_t launch_has_alpha(const uint8_t *d_rgba, uint64_t texels, uint32_t *d_flag, hipStream_t stream) {
    const uint64_t quads = texels / 4;
    const unsigned grid = (unsigned)std::min<uint64_t>(std::max<uint64_t>((quads + 255) / 256, 1), 4096);
    hipLaunchKernelGGL(dev::k_has_alpha, dim3(grid), dim3(256), 0, stream, d_rgba, texels, d_flag);
    return hipGetLastError();
}

hipError_t launch_swizzle_bc3nm(uint8_t *d_rgba, uint64_t texels, hipStream_t stream) {
    const unsigned grid = (unsigned)std::min<uint64_t>(std::max<uint64_t>((texels + 255) / 256, 1), 8192);
    hipLaunchKernelGGL(dev::k_swizzle_bc3nm, dim3(grid), dim3(256), 0, stream, d_rgba, texels);
    return hipGetLastError();
}

hipError_t launch_mips(const MipPlan &p, const uint8_t *d_base, uint8_t *d_ws, hipStream_t stream) {
    if (p.levels <= 1) return hipSuccess;
    const uint32_t nc = p.coef0[p.levels];  // waves
    hipLaunchKernelGGL(dev::k_mip_coeffs, dim3((nc + 3) / 4), dim3(256), 0, stream, p,
                       reinterpret_cast<int32_t *>(d_ws));
    if (p.h0[p.levels]) {
        if (p.w <= kRowMaxWidth) {
            // the row, and with box sums their two buffers (p.w and p.w / 2 words)
            const size_t lds = (size_t)p.w * 4 + (p.boxes_x ? (size_t)p.w * 6 : 0);
            hipLaunchKernelGGL(dev::k_mip_hrows, dim3((unsigned)std::min(p.h, kRowGrid)), dim3(256), lds, stream, p,
                               d_base, d_ws);
        } else {
            hipLaunchKernelGGL(dev::k_mip_pass<false>, dim3(p.h0[p.levels]), dim3(256), 0, stream, p, d_base, d_ws);
        }
    }
    if (p.v0[p.levels])
        hipLaunchKernelGGL(dev::k_mip_pass<true>, dim3(p.v0[p.levels]), dim3(256), 0, stream, p, d_base, d_ws);
    return hipGetLastError();
}

}  // namespace qtx
