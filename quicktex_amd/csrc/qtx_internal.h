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

// Built once per process on the host; uploaded per context.
const DeviceTables &host_tables();

}  // namespace qtx
