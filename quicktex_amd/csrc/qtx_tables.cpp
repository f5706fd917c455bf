// qtx_tables.cpp -- host construction of the codec's constant tables.
//
// Built once per process, uploaded once per context, then read-only on the
// device.  Everything except the trained ordering data (order_tables.inc: the
// ordering lists and the single-bin starting orderings) is
// derived here:
//   * single-colour match tables, brute force per interpolator
//     (quicktex/s3tc/bc1/SingleColorTable.h:45-91);
//   * ordering hashes, least-squares factors and cumulative bin levels
//     (quicktex/s3tc/bc1/OrderTable.h:57-113), with the FP32 operations of the
//     reference in the same order (compiled with -ffp-contract=off);
//   * PreciseRound565 rounding points (quicktex/Color.h:87-95).
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>

#include "qtx_internal.h"

namespace qtx {
namespace {

#include "order_tables.inc"

int s5to8(int v) { return (v << 3) | (v >> 2); }
int s6to8(int v) { return (v << 2) | (v >> 4); }

// Interpolator.cpp:71-138: 2/3 and 1/2 points per interpolator.
int lerp_third(int interp, int a, int b) {
    if (interp == QTX_INTERP_IDEAL_ROUND) return (2 * a + b + 1) / 3;
    if (interp == QTX_INTERP_AMD) return (43 * a + 21 * b + 32) >> 6;
    return (2 * a + b) / 3;
}
int nv_third5(int a, int b) { return (int)((unsigned)((2 * a + b) * 22) / 8u) & 0xFF; }
int nv_third6(int a, int b) {
    const int d = b - a;
    return ((256 * a + d / 4 + 128 + d * 80) >> 8) & 0xFF;
}

void build_match(int interp, int bits, uint32_t *out) {
    const int n = 1 << bits;
    const bool ideal = interp == QTX_INTERP_IDEAL || interp == QTX_INTERP_IDEAL_ROUND;
    for (int target = 0; target < 256; target++) {
        unsigned best = 256;
        int bl = 0, bh = 0;
        for (int lo = 0; lo < n; lo++) {
            const int lo8 = bits == 5 ? s5to8(lo) : s6to8(lo);
            for (int hi = 0; hi < n; hi++) {
                const int hi8 = bits == 5 ? s5to8(hi) : s6to8(hi);
                int v;
                if (interp == QTX_INTERP_NVIDIA) v = bits == 5 ? nv_third5(hi, lo) : nv_third6(hi, lo);
                else v = lerp_third(interp, hi8, lo8) & 0xFF;
                unsigned e = (unsigned)std::abs(v - target);
                if (ideal) e += ((unsigned)std::abs(hi8 - lo8) * 3u) / 100u;
                if (e < best || (e == best && lo == hi)) {
                    best = e;
                    bl = lo;
                    bh = hi;
                }
            }
        }
        out[target] = (uint32_t)bh | ((uint32_t)bl << 8) | (best << 16);
    }
}

void build_orders(int n, const unsigned char *hist, int count, float (*factors)[4], uint32_t *levels,
                  uint16_t *hashes) {
    // LS normal-equation weights for selector s on a d-step line (OrderTable.cpp:40-41)
    const int d = n - 1;
    const float denom = (float)d;
    for (int i = 0; i < count; i++) {
        const unsigned char *h = hist + i * n;
        bool full = false;
        for (int s = 0; s < n; s++) full |= h[s] == 16;
        unsigned packed = 0;
        for (int s = 0; s < n - 1; s++) packed |= (unsigned)h[s] << (4 * s);
        if (!full) hashes[packed] = (uint16_t)i;
        float m[4] = {0.f, 0.f, 0.f, 0.f};
        for (int s = 0; s < n; s++) {
            const float w[4] = {(float)(s * s), (float)(s * (d - s)), (float)(s * (d - s)), (float)((d - s) * (d - s))};
            for (int k = 0; k < 4; k++) m[k] = m[k] + w[k] * (float)h[s];
        }
        const float det = m[0] * m[3] - m[1] * m[2];
        if (std::fabs(det) < 1e-8f) {
            for (int k = 0; k < 4; k++) factors[i][k] = 0.f;
        } else {
            const float scale = (denom / 255.0f) / det;
            factors[i][0] = (m[3] * 1.0f) * scale;
            factors[i][1] = (m[1] * -1.0f) * scale;
            factors[i][2] = (m[2] * -1.0f) * scale;
            factors[i][3] = (m[0] * 1.0f) * scale;
        }
        unsigned l1 = h[0], l2 = l1 + h[1], l3 = n == 4 ? l2 + h[2] : 0;
        levels[i] = l1 | (l2 << 8) | (l3 << 16) | ((full ? 1u : 0u) << 24);
    }
}

void build_midpoints(float *mid, int bits) {
    // Mean of two adjacent expanded codes over 255, as 6-decimal literals; last = 1e37.
    const int n = 1 << bits;
    char buf[32];
    for (int i = 0; i + 1 < n; i++) {
        const int a = bits == 5 ? s5to8(i) : s6to8(i), b = bits == 5 ? s5to8(i + 1) : s6to8(i + 1);
        std::snprintf(buf, sizeof buf, "%.6f", (a + b) / 510.0);
        mid[i] = std::strtof(buf, nullptr);
    }
    mid[n - 1] = 1e37f;
}

DeviceTables *g_tables = nullptr;
std::once_flag g_once;

}  // namespace

const DeviceTables &host_tables() {
    std::call_once(g_once, [] {
        auto *t = new DeviceTables;
        std::memset(t, 0, sizeof *t);
        for (int i = 0; i < 4; i++) {
            build_match(i, 5, t->match[i][0]);
            build_match(i, 6, t->match[i][1]);
        }
        build_orders(4, qtx_orders4, kOrders4, t->factors4, t->levels4, t->hashes4);
        build_orders(3, qtx_orders3, kOrders3, t->factors3, t->levels3, t->hashes3);
        // starting ordering of a single-bin histogram, by bin (OrderTable.cpp:43-44; the
        // 3-colour list maps bins 1 and 2 crosswise and is kept as the reference has it)
        for (int s = 0; s < 4; s++) t->single4[s] = qtx_single_color_hashes4[s];
        for (int s = 0; s < 3; s++) t->single3[s] = qtx_single_color_hashes3[s];
        std::memcpy(t->best4, qtx_best_orders4, sizeof t->best4);
        std::memcpy(t->best3, qtx_best_orders3, sizeof t->best3);
        build_midpoints(t->mid5, 5);
        build_midpoints(t->mid6, 6);
        g_tables = t;
    });
    return *g_tables;
}

}  // namespace qtx
