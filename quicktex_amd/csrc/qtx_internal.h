// qtx_internal.h -- definitions shared by the host runtime and the gfx950 kernels.
#pragma once
#include <cstdint>

#include "../../include/qtx.h"

namespace qtx {

constexpr int kOrders4 = 969;      // compositions of 16 into 4 bins (OrderTable.h:37)
constexpr int kOrders3 = 153;      // compositions of 16 into 3 bins
constexpr int kBestOrders4 = 128;  // OrderTable.h:42 (RGBCX_USE_SMALLER_TABLES unset)
constexpr int kBestOrders3 = 32;

// Device-resident constant tables, built once per context (host: qtx_tables.cpp).
struct DeviceTables {
    // Single-colour match tables per interpolator and bit depth (SingleColorTable.h:45-91):
    // entry = high | low << 8 | error << 16.
    uint32_t match[4][2][256];
    // Cluster-fit orderings.  factors: the LS inverse per ordering (OrderTable.h:63-82);
    // levels: cumulative histogram bins (h0, h0+h1, h0+h1+h2) | single-colour flag << 24.
    float factors4[kOrders4][4];
    float factors3[kOrders3][4];
    uint32_t levels4[kOrders4];
    uint32_t levels3[kOrders3];
    uint16_t hashes4[4096];  // packed histogram -> ordering index
    uint16_t hashes3[256];
    uint16_t single4[4];  // ordering index of the histogram with all 16 in bin i
    uint16_t single3[4];
    uint16_t best4[kOrders4 * kBestOrders4];
    uint16_t best3[kOrders3 * kBestOrders3];
    // PreciseRound565 rounding points (Color.h:87-95)
    float mid5[32];
    float mid6[64];
};

// Kernel-side encoder settings: qtx_bc1_params after clamping.
struct BC1Settings {
    int32_t color_mode;
    int32_t error_mode;
    int32_t endpoint_mode;
    int32_t ls_passes;  // 1 or 2
    int32_t ep_passes;  // 1 or 2 (already folded with needs_block_error)
    int32_t cf_passes;  // 1 or 2
    int32_t exhaustive;
    uint32_t power_iterations;
    uint32_t search_rounds;
    uint32_t orderings4;
    uint32_t orderings3;
};

// One texture of a batch as the batch kernels see it (device-resident array,
// sorted so that the work units of job k are [unit_begin, next job's unit_begin)).
enum BatchKind : int32_t {
    BJ_ENC_BC1 = 0x01, BJ_ENC_BC3 = 0x03, BJ_ENC_BC4 = 0x04, BJ_ENC_BC5 = 0x05,
    BJ_DEC_BC1 = 0x11, BJ_DEC_BC3 = 0x13, BJ_DEC_BC4 = 0x14, BJ_DEC_BC5 = 0x15,
};
struct BatchJob {
    const uint8_t *src;  // RGBA8 texture (encode) or blocks (decode)
    uint8_t *dst;        // blocks (encode) or RGBA8 texture (decode)
    int32_t w, h, bw, bh;
    uint32_t nblocks;      // bw * bh
    uint32_t unit_begin;   // first work unit of this job within its kernel's queue
    uint32_t unit_blocks;  // blocks per work unit (a multiple of 4 for colour jobs)
    int32_t kind;          // BatchKind
    int32_t interp;        // BC1/BC3 interpolator
    int32_t p0, p1;        // BC4: channel; BC5: channels; BC1 decode: write_alpha
    BC1Settings s;         // BC1/BC3 encode
};

// Built once per process on the host; uploaded per context.
const DeviceTables &host_tables();

}  // namespace qtx
