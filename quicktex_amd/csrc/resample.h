// resample.h -- the mip chain of quicktex.dds.encode on the GPU (host plan and
// launchers; kernels in resample.hip).
//
// quicktex.dds.encode (quicktex/dds.py:276-287) makes every mip level from the
// base image with image_utils.resize_no_premultiply (quicktex/image_utils.py:42-52):
// Pillow's Image.resize with Image.BILINEAR, applied to the RGB bands and to the
// alpha band separately.  Pillow resamples each band independently with the same
// coefficients, so the level is that resample applied to every channel of the
// RGBA8 base.  Pillow's algorithm (libImaging/Resample.c; Pillow is an unpinned
// dependency of the reference, pyproject.toml:30):
//   * per axis and output index i, a triangle filter of support `scale`
//     (= in / out when shrinking, else 1) centred at (i + 0.5) * scale; its window
//     [lo, lo + n) is clipped to the image, the weights are normalised to sum 1 in
//     double precision and stored as fixed point with 22 fraction bits, rounded
//     half away from zero;
//   * a horizontal pass into an 8-bit image, then a vertical pass; each value is
//     clamp((2^21 + sum(v * k)) >> 22, 0, 255); a pass is skipped when its axis
//     keeps its size.
// The coefficients are computed on the device, one output index per thread, with
// the same double operations in the same order (contraction off), so they are the
// bit patterns Pillow computes.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace qtx {

constexpr int kMaxMipLevels = 32;  // a 2^31-texel texture has at most 32 levels
constexpr int kResampleWideTaps = 64;  // windows longer than this use a wave per output
constexpr int kRowMaxWidth = 8192;     // widest base whose rows the horizontal pass keeps in LDS
constexpr int kRowGrid = 1024;         // workgroups of that pass (they stride over the rows)
constexpr int kBoxMaxWidth = 4096;     // widest base whose rows also get box sums in LDS (k_mip_hrows)

// Layout of one mip chain in a caller's workspace, and its launch ranges.  Passed
// to the kernels by value.
struct MipPlan {
    int32_t w, h, levels;  // base size; levels of the chain, the base included
    int32_t lw[kMaxMipLevels], lh[kMaxMipLevels];    // level sizes
    int32_t ksx[kMaxMipLevels], ksy[kMaxMipLevels];  // coefficient slots per output index
    uint64_t cx[kMaxMipLevels], cy[kMaxMipLevels];   // workspace offsets (int32 words) of the
    uint64_t wx[kMaxMipLevels], wy[kMaxMipLevels];   //   coefficient and window (lo, n) tables
    uint64_t tmp[kMaxMipLevels];                     // byte offset of the horizontal-pass image
    // first wave / unit of each level in the coefficient, horizontal and vertical
    // launches (prefix sums; entry `levels` is the total)
    uint32_t coef0[kMaxMipLevels + 1];
    uint32_t h0[kMaxMipLevels + 1], v0[kMaxMipLevels + 1];
    int32_t gx[kMaxMipLevels];        // lanes per output in the horizontal pass over LDS rows
    uint32_t exact_x, exact_y;        // bit m: that axis of level m shrinks by a power of two
    uint32_t boxes_x;                 // bit m: level m's interior outputs come from box sums (k_mip_hrows)
    uint8_t *dst[kMaxMipLevels];  // level outputs, RGBA8 (dst[0] unused)
};

// Sizes of quicktex.image_utils.mip_sizes (quicktex/image_utils.py:9-38): level 0
// is (w, h), each next level halves both sides (rounding down, at least 1); the
// chain ends at 1x1 or after `requested` levels (requested <= 0: the full chain).
int mip_level_count(int32_t w, int32_t h, int32_t requested);
// Fills the plan for a chain of `levels` levels; returns the workspace bytes it needs.
size_t mip_plan(int32_t w, int32_t h, int32_t levels, MipPlan &p);
// Enqueues the coefficient, horizontal and vertical launches (dst pointers set).
hipError_t launch_mips(const MipPlan &p, const uint8_t *d_base, uint8_t *d_ws, hipStream_t stream);

// Sets *d_flag to 1 if any of the texels has alpha below 255 (the flag must be 0
// before; the kernel only ever writes 1).  d_rgba 16-byte aligned.
hipError_t launch_has_alpha(const uint8_t *d_rgba, uint64_t texels, uint32_t *d_flag, hipStream_t stream);
// (R, G, B, A) -> (255, G, G, R) in place.
hipError_t launch_swizzle_bc3nm(uint8_t *d_rgba, uint64_t texels, hipStream_t stream);

}  // namespace qtx
