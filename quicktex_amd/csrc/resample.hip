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
// filter): the same double operations in the same order as Pillow.
__device__ void axis_coeffs(int in_size, int out_size, int i, int32_t *__restrict__ k, int32_t *__restrict__ win) {
    const double scale = (double)in_size / (double)out_size;
    const double filterscale = scale < 1.0 ? 1.0 : scale;
    const double support = 1.0 * filterscale;
    const double center = 0.0 + (i + 0.5) * scale;
    const double ss = 1.0 / filterscale;
    int xmin = (int)(center - support + 0.5);
    if (xmin < 0) xmin = 0;
    int xmax = (int)(center + support + 0.5);
    if (xmax > in_size) xmax = in_size;
    xmax -= xmin;
    double ww = 0.0;
    for (int x = 0; x < xmax; x++) {
        double t = (x + xmin - center + 0.5) * ss;
        if (t < 0.0) t = -t;
        ww += t < 1.0 ? 1.0 - t : 0.0;
    }
    for (int x = 0; x < xmax; x++) {
        double t = (x + xmin - center + 0.5) * ss;
        if (t < 0.0) t = -t;
        double v = t < 1.0 ? 1.0 - t : 0.0;
        if (ww != 0.0) v /= ww;
        k[x] = v < 0.0 ? (int32_t)(-0.5 + v * (1 << kBits)) : (int32_t)(0.5 + v * (1 << kBits));
    }
    win[0] = xmin;
    win[1] = xmax;
}

// Level of a launch-wide index: the last level whose first index is <= u.
__device__ __forceinline__ int level_of(const uint32_t *first, int levels, uint32_t u) {
    int m = 0;
    for (int l = 1; l < levels; l++)
        if (first[l] <= u) m = l;
    return m;
}

// One thread per output index of every level and axis: x indices, then y.
__global__ __launch_bounds__(256) void k_mip_coeffs(MipPlan p, int32_t *__restrict__ ws) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= p.coef0[p.levels]) return;
    const int m = level_of(p.coef0, p.levels, t);
    const int i = (int)(t - p.coef0[m]);
    if (i < p.lw[m]) axis_coeffs(p.w, p.lw[m], i, ws + p.cx[m] + (size_t)i * p.ksx[m], ws + p.wx[m] + 2 * i);
    else axis_coeffs(p.h, p.lh[m], i - p.lw[m], ws + p.cy[m] + (size_t)(i - p.lw[m]) * p.ksy[m],
                     ws + p.wy[m] + 2 * (i - p.lw[m]));
}

__device__ __forceinline__ void accumulate(uint32_t px, int32_t k, int32_t (&ss)[4]) {
#pragma unroll
    for (int c = 0; c < 4; c++) ss[c] += (int32_t)((px >> (8 * c)) & 0xFFu) * k;
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
    const int32_t *k = kt + (size_t)i * ks;
    int32_t ss[4] = {0, 0, 0, 0};
    const int t0 = wide ? (int)(threadIdx.x & 63) : 0, dt = wide ? 64 : 1;
    for (int t = t0; t < n; t += dt) {
        const uint32_t px = VERTICAL ? src[(size_t)(lo + t) * sw + x] : src[(size_t)y * sw + lo + t];
        accumulate(px, k[t], ss);
    }
    if (wide) wave_sum(ss);
    if (!wide || (threadIdx.x & 63) == 0) {
#pragma unroll
        for (int c = 0; c < 4; c++) ss[c] += 1 << (kBits - 1);
        dst[(size_t)y * lw + x] = pack_clip(ss);
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
        coef += (uint32_t)(p.lw[m] + p.lh[m]);
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

hipError_t launch_mips(const MipPlan &p, const uint8_t *d_base, uint8_t *d_ws, hipStream_t stream) {
    if (p.levels <= 1) return hipSuccess;
    const uint32_t nc = p.coef0[p.levels];
    hipLaunchKernelGGL(dev::k_mip_coeffs, dim3((nc + 255) / 256), dim3(256), 0, stream, p,
                       reinterpret_cast<int32_t *>(d_ws));
    if (p.h0[p.levels])
        hipLaunchKernelGGL(dev::k_mip_pass<false>, dim3(p.h0[p.levels]), dim3(256), 0, stream, p, d_base, d_ws);
    if (p.v0[p.levels])
        hipLaunchKernelGGL(dev::k_mip_pass<true>, dim3(p.v0[p.levels]), dim3(256), 0, stream, p, d_base, d_ws);
    return hipGetLastError();
}

}  // namespace qtx
