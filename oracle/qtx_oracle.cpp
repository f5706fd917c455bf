// oracle/qtx_oracle.cpp -- CPU restatement of the reference S3TC path.
//
// TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and
// bench.py's cpu_baseline leg load liboracle.so, as the parity checker (or as
// the CPU baseline when the compiled reference is absent).  The product
// (quicktex_amd/) never links or calls it, and has no CPU fallback.
//
// It restates, function by function, the reference quicktex 0.3.1 codecs
// (citations are path:line in the reference tree), keeping every FP32
// operation in the reference's order and contraction off (-ffp-contract=off),
// and every tie rule as written.
//
// Parity pins (tests/test_oracle.py): the reference test goldens
// (tests/images/__init__.py:18-20,31-32; test_bc1.py:9-12; test_bc4.py:9-11,
// 58-70), the compiled reference in oracle/_ref (random blocks, every level,
// colour mode and interpolator), and closed forms / brute force (BC4
// optimality, single-colour tables, order-table structure).
//
// Known, pinned deviation: EndpointMode::LeastSquares reads an uninitialised
// accumulator in the reference (quicktex/s3tc/bc1/BC1Encoder.cpp:510).  This
// restatement zero-initialises it; see the note at find_endpoints().

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <vector>

#include "../include/qtx.h"
#include "../quicktex_amd/csrc/order_tables.inc"

namespace {

// ----------------------------------------------------------------------------
// Algorithmic work accounting.  Each primitive of the encoder counts its
// invocations; qtxo_op_weights() gives the static number of 32-bit lane
// operations one invocation needs on a GPU lane (16 texels, the reference's
// formulation), counting a dot product of two RGB byte triples as one operation
// (the hardware has it as one instruction, v_dot4_u32_u8) and every other add,
// multiply, compare, select, conversion, shift or table read as one.  A selector
// search is priced by what its error needs (the selectors of a trial that is not
// kept are not work the algorithm requires).  The full selector searches, which
// dominate the costly levels, are priced at the fewest gfx950 vector
// instructions that compute them: a palette read as one byte permute per channel
// and pair from per-channel tables (6), and per texel one dot product and one
// shift-add per candidate, the maximum over the candidates with 3-input maxima
// and the running sum with 3-input adds (10 for 4 candidates, 7 for 3), so that
// no encoder can run them in fewer instructions than they are charged.
// ops = sum(count * weight) is the "algorithmic ops" of a workload.  Counts are
// per thread: count single-threaded.
enum OpKind {
    OP_METRICS, OP_SOLID, OP_EP_GREY, OP_EP_LS, OP_EP_BB, OP_EP_BBINT, OP_EP_PCA, OP_PCA_ITER,
    OP_SEL_FULL4, OP_SEL_FULL3, OP_SEL_CHECK2, OP_SEL_FASTER, OP_SEL_NONE,
    OP_LS_PX, OP_LS_SUMS, OP_SC_PX, OP_CF_SETUP, OP_SEARCH_ROUND, OP_WRITE, OP_BC4, OP_KINDS
};
const uint32_t kOpWeight[OP_KINDS] = {
    290,  // metrics: 16 x (3 extracts, 3 min, 3 max, 3 adds, grey and black tests) + 3 averages
    30,   // solid block: 3 table lookups, packing, order fix-up
    15,   // greyscale endpoints
    150,  // least-squares endpoints: 16 x (3 extracts, 3 mul-adds) + solve + 2 roundings
    130,  // bounding box (float): insets, 16 x covariance sign terms, 2 roundings
    115,  // bounding box (int)
    360,  // PCA: covariance 16 x 9, scaling, 16 projections x 11
    32,   // one PCA power iteration: 4 rows x (4 mul + 4 add)
    166,  // 4-candidate full selector search: palette 6 + 16 x (4 dots, 4 shift-adds, 2 max3; sum)
    118,  // 3-candidate full selector search: palette 6 + 16 x (3 dots, 3 shift-adds, 1 max3; sum)
    338,  // Check2 selector search: set-up 50 + 16 x 18
    306,  // Faster selector search: set-up 50 + 16 x 16
    210,  // selector search without error: set-up 50 + 16 x 10
    200,  // least squares from texels: 16 x 8 + solve + 2 roundings
    80,   // least squares from prefix sums: 3 reads, sums, 2 x 3 mul-adds, 2 roundings
    85,   // single-colour fallback error: palette + 16 dots + sums + 3 table reads
    430,  // cluster-fit set-up: histogram, 16 projections, 63-comparator sort, prefix sums
    8,    // one endpoint-search move
    42,   // block packing
    180,  // one BC4 channel block
};
thread_local uint64_t g_ops[OP_KINDS];
inline void count(OpKind k, uint64_t n = 1) { g_ops[k] += n; }

// ----------------------------------------------------------------------------
// Colour helpers: quicktex/Color.h:29-96, quicktex/Color.cpp, quicktex/util.h

struct C4 {
    uint8_t v[4];  // r, g, b, a
};
inline C4 mk(int r, int g, int b, int a = 255) { return C4{{(uint8_t)r, (uint8_t)g, (uint8_t)b, (uint8_t)a}}; }
inline bool eq(const C4 &x, const C4 &y) { return std::memcmp(x.v, y.v, 4) == 0; }

// util.h:110-126
inline int scale8to5(int v) { int v2 = v * 31 + 128; return (v2 + (v2 >> 8)) >> 8; }
inline int scale8to6(int v) { int v2 = v * 63 + 128; return (v2 + (v2 >> 8)) >> 8; }
inline int scale5to8(int v) { return (v << 3) | (v >> 2); }
inline int scale6to8(int v) { return (v << 2) | (v >> 4); }
// util.h:145
inline uint8_t clamp255(int32_t i) { return (uint8_t)(((unsigned)i & 0xFFFFFF00U) ? (~(i >> 31)) : i); }
inline int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }
// Color.cpp:138-143
inline uint16_t pack565u(const C4 &c) { return (uint16_t)(c.v[2] | (c.v[1] << 5) | (c.v[0] << 11)); }
// Color.cpp:218-219
inline C4 scale_to565(const C4 &c) { return mk(scale8to5(c.v[0]), scale8to6(c.v[1]), scale8to5(c.v[2])); }
inline C4 scale_from565(const C4 &c) { return mk(scale5to8(c.v[0]), scale6to8(c.v[1]), scale5to8(c.v[2])); }
// Color.h:81-82
inline bool is_black(const C4 &c) { return (c.v[0] | c.v[1] | c.v[2]) < 4; }

// Color.h:87-95: rounding points between adjacent 5/6-bit codes, i.e. the mean of
// the two expanded 8-bit values over 255, printed to 6 decimals; the last is 1e37.
float g_mid5[32], g_mid6[64];
void init_midpoints() {
    char buf[32];
    for (int i = 0; i < 31; i++) {
        std::snprintf(buf, sizeof buf, "%.6f", (scale5to8(i) + scale5to8(i + 1)) / 2.0 / 255.0);
        g_mid5[i] = std::strtof(buf, nullptr);
    }
    g_mid5[31] = 1e37f;
    for (int i = 0; i < 63; i++) {
        std::snprintf(buf, sizeof buf, "%.6f", (scale6to8(i) + scale6to8(i + 1)) / 2.0 / 255.0);
        g_mid6[i] = std::strtof(buf, nullptr);
    }
    g_mid6[63] = 1e37f;
}

// Color.cpp:163-183
C4 precise_round565(const float v[3]) {
    int tr = (int)(v[0] * 31.0f);
    int tg = (int)(v[1] * 63.0f);
    int tb = (int)(v[2] * 31.0f);
    int r = clampi(tr, 0, 31), g = clampi(tg, 0, 63), b = clampi(tb, 0, 31);
    r += v[0] > g_mid5[r];
    g += v[1] > g_mid6[g];
    b += v[2] > g_mid5[b];
    return mk(r, g, b);
}

// ----------------------------------------------------------------------------
// Interpolators: quicktex/s3tc/interpolator/Interpolator.cpp:47-138

int interp8(int t, int v0, int v1) {
    switch (t) {
        case QTX_INTERP_IDEAL_ROUND: return (v0 * 2 + v1 + 1) / 3;          // :75
        case QTX_INTERP_AMD: return (v0 * 43 + v1 * 21 + 32) >> 6;          // :136
        default: return (v0 * 2 + v1) / 3;                                  // :71 (Nvidia inherits)
    }
}
int half8(int t, int v0, int v1) {
    if (t == QTX_INTERP_AMD) return (v0 + v1 + 1) >> 1;  // :138
    return (v0 + v1) / 2;                                // :73
}
int interp5(int t, int v0, int v1) {
    if (t == QTX_INTERP_NVIDIA) return (uint8_t)(((2 * v0 + v1) * 22) / 8U);  // :88-91
    return interp8(t, scale5to8(v0), scale5to8(v1));
}
int interp6(int t, int v0, int v1) {
    if (t == QTX_INTERP_NVIDIA) {  // :93-97
        const int gd = v1 - v0;
        return (uint8_t)((256 * v0 + (gd / 4) + 128 + gd * 80) >> 8);
    }
    return interp8(t, scale6to8(v0), scale6to8(v1));
}
int half5(int t, int v0, int v1) {
    if (t == QTX_INTERP_NVIDIA) return (uint8_t)(((v0 + v1) * 33) / 8U);  // :99-102
    return half8(t, scale5to8(v0), scale5to8(v1));
}
int half6(int t, int v0, int v1) {
    if (t == QTX_INTERP_NVIDIA) {  // :104-108
        const int gd = v1 - v0;
        return (uint8_t)((256 * v0 + gd / 4 + 128 + gd * 128) >> 8);
    }
    return half8(t, scale6to8(v0), scale6to8(v1));
}

// InterpolateBC1: Interpolator.cpp:57-73 and the Nvidia override :110-127.
// low/high are unscaled 5:6:5 colours.
void interpolate_bc1(int t, const C4 &low, const C4 &high, bool use_3color, C4 out[4]) {
    out[0] = scale_from565(low);
    out[1] = scale_from565(high);
    if (t == QTX_INTERP_NVIDIA) {
        if (!use_3color) {
            out[2] = mk(interp5(t, low.v[0], high.v[0]), interp6(t, low.v[1], high.v[1]), interp5(t, low.v[2], high.v[2]));
            out[3] = mk(interp5(t, high.v[0], low.v[0]), interp6(t, high.v[1], low.v[1]), interp5(t, high.v[2], low.v[2]));
        } else {
            out[2] = mk(half5(t, low.v[0], high.v[0]), half6(t, low.v[1], high.v[1]), half5(t, low.v[2], high.v[2]));
            out[3] = mk(0, 0, 0, 0);
        }
        return;
    }
    if (use_3color) {
        out[2] = mk(half8(t, out[0].v[0], out[1].v[0]), half8(t, out[0].v[1], out[1].v[1]), half8(t, out[0].v[2], out[1].v[2]));
        out[3] = mk(0, 0, 0, 0);
    } else {
        out[2] = mk(interp8(t, out[0].v[0], out[1].v[0]), interp8(t, out[0].v[1], out[1].v[1]), interp8(t, out[0].v[2], out[1].v[2]));
        out[3] = mk(interp8(t, out[1].v[0], out[0].v[0]), interp8(t, out[1].v[1], out[0].v[1]), interp8(t, out[1].v[2], out[0].v[2]));
    }
}

// ----------------------------------------------------------------------------
// Single-colour match tables: quicktex/s3tc/bc1/SingleColorTable.h:45-91.
// The colour-count template argument is unused by the reference, so the
// "3-colour" tables equal the 4-colour ones; one table per (interp, bits).

struct Match {
    uint8_t high, low, error;
};
Match g_match[4][2][256];  // [interp][bits==6][value]

void build_match_table(int t, int bits, Match *m) {
    const int size = 1 << bits;
    const bool ideal = (t == QTX_INTERP_IDEAL || t == QTX_INTERP_IDEAL_ROUND);
    const bool use_8bit = (t != QTX_INTERP_NVIDIA);
    for (int i = 0; i < 256; i++) {
        unsigned error = 256;
        for (int low = 0; low < size; low++) {
            const int low8 = bits == 5 ? scale5to8(low) : scale6to8(low);
            for (int high = 0; high < size; high++) {
                const int high8 = bits == 5 ? scale5to8(high) : scale6to8(high);
                int value;
                if (use_8bit) value = (uint8_t)interp8(t, high8, low8);
                else value = bits == 5 ? interp5(t, high, low) : interp6(t, high, low);
                unsigned new_error = (unsigned)std::abs(value - i);
                if (ideal) new_error += ((unsigned)std::abs(high8 - low8) * 3) / 100;
                if (new_error < error || (new_error == error && low == high)) {
                    m[i].low = (uint8_t)low;
                    m[i].high = (uint8_t)high;
                    m[i].error = (uint8_t)new_error;
                    error = new_error;
                }
            }
        }
    }
}

// ----------------------------------------------------------------------------
// Ordering tables: quicktex/s3tc/bc1/OrderTable.h:57-113, Histogram.h:70-79.
// Weights[sel] = (sel^2, sel*(d-sel), sel*(d-sel), (d-sel)^2), d = colours-1
// (OrderTable.cpp:40-41).

struct Orders {
    int n;               // 3 or 4 colours
    int count;           // 153 or 969
    int best_count;      // 32 or 128
    const uint8_t *hist;   // count x n
    const uint16_t *best;  // count x best_count
    uint16_t hashes[4096];
    uint16_t single_color_hash[4];
    float factors[969][4];
    float weights[4][4];
};
Orders g_o3, g_o4;

unsigned packed_hist(const uint8_t *h, int n) {
    unsigned p = 0;
    for (int i = 0; i < n - 1; i++) p |= (unsigned)h[i] << (i * 4);
    return p;
}

void build_orders(Orders &o, int n) {
    o.n = n;
    o.count = n == 4 ? 969 : 153;
    o.best_count = n == 4 ? 128 : 32;
    o.hist = n == 4 ? qtx_orders4 : qtx_orders3;
    o.best = n == 4 ? qtx_best_orders4 : qtx_best_orders3;
    std::memset(o.hashes, 0, sizeof o.hashes);
    const int d = n - 1;
    for (int s = 0; s < n; s++) {
        o.weights[s][0] = (float)(s * s);
        o.weights[s][1] = (float)(s * (d - s));
        o.weights[s][2] = (float)(s * (d - s));
        o.weights[s][3] = (float)((d - s) * (d - s));
    }
    // SingleColorHashes (OrderTable.cpp:43-44) are data: the 3-colour list maps
    // bins 1 and 2 crosswise, and GetHash (OrderTable.h:91-104) indexes it by bin.
    for (int s = 0; s < n; s++) o.single_color_hash[s] = n == 4 ? qtx_single_color_hashes4[s] : qtx_single_color_hashes3[s];
    const float denominator = n == 4 ? 3.0f : 2.0f;
    for (int i = 0; i < o.count; i++) {
        const uint8_t *h = o.hist + i * n;
        bool any16 = false;
        for (int s = 0; s < n; s++) any16 |= h[s] == 16;
        if (!any16) o.hashes[packed_hist(h, n)] = (uint16_t)i;
        float m[4] = {0, 0, 0, 0};
        for (int s = 0; s < n; s++)
            for (int k = 0; k < 4; k++) m[k] = m[k] + o.weights[s][k] * (float)h[s];
        float det = (m[0] * m[3]) - (m[1] * m[2]);
        if (std::fabs(det) < 1e-8f) {
            o.factors[i][0] = o.factors[i][1] = o.factors[i][2] = o.factors[i][3] = 0;
        } else {
            std::swap(m[0], m[3]);
            m[0] = m[0] * 1.0f; m[1] = m[1] * -1.0f; m[2] = m[2] * -1.0f; m[3] = m[3] * 1.0f;
            const float s = (denominator / 255.0f) / det;
            for (int k = 0; k < 4; k++) o.factors[i][k] = m[k] * s;
        }
    }
}

uint16_t get_hash(const Orders &o, const uint8_t *h) {
    for (int s = 0; s < o.n; s++)
        if (h[s] == 16) return o.single_color_hash[s];
    return o.hashes[packed_hist(h, o.n)];
}
bool is_single_color_hash(const Orders &o, uint16_t hash) {
    for (int s = 0; s < o.n; s++)
        if (o.single_color_hash[s] == hash) return true;
    return false;
}

std::once_flag g_init_once;
void init_tables() {
    std::call_once(g_init_once, [] {
        init_midpoints();
        for (int t = 0; t < 4; t++) {
            build_match_table(t, 5, g_match[t][0]);
            build_match_table(t, 6, g_match[t][1]);
        }
        build_orders(g_o3, 3);
        build_orders(g_o4, 4);
    });
}

// ----------------------------------------------------------------------------
// Block metrics: quicktex/ColorBlock.h:80-117

struct Block {
    C4 px[16];  // index x + 4y
};

struct Metrics {
    C4 min, max, avg;
    bool is_greyscale, has_black;
    int sums[3];
};

bool is_single_color(const Block &b) {
    for (int i = 1; i < 16; i++)
        if (!eq(b.px[i], b.px[0])) return false;
    return true;
}

Metrics get_metrics(const Block &b, bool ignore_black) {
    count(OP_METRICS);
    Metrics m;
    m.min = mk(255, 255, 255);
    m.max = mk(0, 0, 0);
    m.avg = mk(0, 0, 0);  // Color() default: (0,0,0,255)
    m.has_black = false;
    m.is_greyscale = true;
    m.sums[0] = m.sums[1] = m.sums[2] = 0;
    unsigned total = 0;
    for (int i = 0; i < 16; i++) {
        const C4 &v = b.px[i];
        const bool blk = is_black(v);
        m.has_black |= blk;
        if (ignore_black && blk) continue;
        m.is_greyscale &= (v.v[0] == v.v[1]) && (v.v[0] == v.v[2]);
        for (int c = 0; c < 3; c++) {
            m.min.v[c] = std::min(m.min.v[c], v.v[c]);
            m.max.v[c] = std::max(m.max.v[c], v.v[c]);
            m.sums[c] += v.v[c];
        }
        total++;
    }
    if (total > 0) {
        for (int c = 0; c < 3; c++) m.avg.v[c] = (uint8_t)((m.sums[c] + (int)(total / 2)) / (int)total);
        m.avg.v[3] = (uint8_t)((int)(total / 2) / (int)total);  // alpha sum is 0
    }
    return m;
}

// ----------------------------------------------------------------------------
// BC1 encoder: quicktex/s3tc/bc1/BC1Encoder.cpp

enum { MODE_INCOMPLETE = 0 };

struct Result {
    C4 low = mk(0, 0, 0), high = mk(0, 0, 0);
    uint8_t sel[16] = {0};
    int mode = MODE_INCOMPLETE;
    unsigned error = UINT_MAX;
};

struct Enc {
    qtx_bc1_params p;
};

inline int dot3(const int a[3], const int b[3]) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

// FindEndpointsSingleColor(result, color, is_3color): BC1Encoder.cpp:443-456
void endpoints_single_color(const Enc &e, Result &r, const C4 &color, bool is_3color) {
    const Match *m5 = g_match[e.p.interpolator][0];
    const Match *m6 = g_match[e.p.interpolator][1];
    const Match &mr = m5[color.v[0]], &mg = m6[color.v[1]], &mb = m5[color.v[2]];
    r.mode = is_3color ? QTX_COLOR_THREE : QTX_COLOR_FOUR;
    r.error = (unsigned)mr.error + mg.error + mb.error;
    r.low = mk(mr.low, mg.low, mb.low);
    r.high = mk(mr.high, mg.high, mb.high);
}

// FindEndpointsSingleColor(result, pixels, color, is_3color): BC1Encoder.cpp:458-471.
// The error is measured against colour 2 of the endpoints the result held on
// entry, with alpha included.
void endpoints_single_color_px(const Enc &e, Result &r, const Block &b, const C4 &color, bool is_3color) {
    count(OP_SC_PX);
    C4 pal[4];
    interpolate_bc1(e.p.interpolator, r.low, r.high, is_3color, pal);
    const C4 ref = pal[2];
    endpoints_single_color(e, r, color, is_3color);
    r.error = 0;
    for (int i = 0; i < 16; i++) {
        unsigned err = 0;
        for (int c = 0; c < 4; c++) {
            int d = (int)b.px[i].v[c] - (int)ref.v[c];
            err += (unsigned)(d * d);
        }
        r.error += err;
        r.sel[i] = 1;
    }
}

// FindEndpoints: BC1Encoder.cpp:473-677
void find_endpoints(const Enc &e, Result &r, const Block &b, const Metrics &m, int endpoint_mode, bool ignore_black) {
    if (m.is_greyscale) {
        count(OP_EP_GREY);
        const int fr = b.px[0].v[0];  // pixel (0,0) even when ignoring black (BC1Encoder.cpp:476)
        if ((int)m.max.v[0] - (int)m.min.v[0] < 2) {
            const int f5 = scale8to5(fr), f6 = scale8to6(fr);
            r.low = mk(f5, f6, f5);
            r.high = r.low;
        } else {
            r.low = mk(scale8to5(m.min.v[0]), scale8to6(m.min.v[0]), scale8to5(m.min.v[0]));
            r.high = mk(scale8to5(m.max.v[0]), scale8to6(m.max.v[0]), scale8to5(m.max.v[0]));
        }
    } else if (endpoint_mode == QTX_EP_LEAST_SQUARES) {
        count(OP_EP_LS);
        // BC1Encoder.cpp:495-552.  The reference leaves sums_xy uninitialised
        // (:510); it is zero here.  Both endpoints are placed at min on the primary
        // axis (:505-506), so low == high whatever the accumulator holds.
        C4 diff = mk(m.max.v[0] - m.min.v[0], m.max.v[1] - m.min.v[1], m.max.v[2] - m.min.v[2]);
        float l[4] = {0, 0, 0, 0}, h[4] = {0, 0, 0, 0};
        unsigned chan0;
        if (diff.v[0] >= diff.v[1] && diff.v[0] >= diff.v[2]) chan0 = 0;  // Color.cpp:197-201
        else if (diff.v[1] >= diff.v[2] && diff.v[1] >= diff.v[0]) chan0 = 1;
        else chan0 = 2;
        l[chan0] = (float)m.min.v[chan0];
        h[chan0] = (float)m.min.v[chan0];
        unsigned sums_xy[3] = {0, 0, 0};
        for (int i = 0; i < 16; i++)
            for (unsigned c = 0; c < 3; c++) sums_xy[c] += (unsigned)b.px[i].v[chan0] * b.px[i].v[c];
        const unsigned sum_x = (unsigned)m.sums[chan0];
        const unsigned sum_xx = sums_xy[chan0];
        const float denominator = (float)(16 * sum_xx) - (float)(sum_x * sum_x);
        if (std::fabs(denominator) > 1e-8f) {
            for (unsigned i = 1; i < 3; i++) {
                const unsigned chan = (chan0 + i) % 3;
                const unsigned sum_y = (unsigned)m.sums[chan];
                const unsigned sum_xy = sums_xy[chan];
                const float a = (float)((16 * sum_xy) - (sum_x * sum_y)) / denominator;
                const float bb = (float)((sum_xx * sum_y) - (sum_xy * sum_x)) / denominator;
                l[chan] = bb + (a * l[chan0]);
                h[chan] = bb + (a * h[chan0]);
            }
        }
        for (int c = 0; c < 3; c++) {
            const float inset = (h[c] - l[c]) / 16.0f;
            l[c] = ((l[c] + inset) / 255.0f);
            h[c] = ((h[c] - inset) / 255.0f);
        }
        r.low = precise_round565(l);
        r.high = precise_round565(h);
    } else if (endpoint_mode == QTX_EP_BOUNDING_BOX) {
        count(OP_EP_BB);
        // BC1Encoder.cpp:553-580 (note the operator precedence at :572-573)
        float l[4] = {0, 0, 0, 0}, h[4] = {0, 0, 0, 0};
        const float bias = 8.0f / 255.0f;
        for (int c = 0; c < 3; c++) {
            l[c] = (float)m.min.v[c] / 255.0f;
            h[c] = (float)m.max.v[c] / 255.0f;
            const float inset = (h[c] - l[c] - bias) / 16.0f;
            l[c] += inset;
            h[c] -= inset;
        }
        int icov_xz = 0, icov_yz = 0;
        for (int i = 0; i < 16; i++) {
            const int bl = (int)b.px[i].v[2] - m.avg.v[2];
            icov_xz += bl * (int)b.px[i].v[0] - m.avg.v[0];
            icov_yz += bl * (int)b.px[i].v[1] - m.avg.v[1];
        }
        if (icov_xz < 0) std::swap(l[0], h[0]);
        if (icov_yz < 0) std::swap(l[1], h[1]);
        r.low = precise_round565(l);
        r.high = precise_round565(h);
    } else if (endpoint_mode == QTX_EP_BOUNDING_BOX_INT) {
        count(OP_EP_BBINT);
        // BC1Encoder.cpp:581-605
        C4 mn = mk(0, 0, 0), mx = mk(0, 0, 0);
        for (int c = 0; c < 3; c++) {
            const int inset = ((int)(m.max.v[c] - m.min.v[c]) - 8) >> 4;
            mn.v[c] = clamp255(m.min.v[c] + inset);
            mx.v[c] = clamp255(m.max.v[c] - inset);
        }
        int icov_xz = 0, icov_yz = 0;
        for (int i = 0; i < 16; i++) {
            const int bl = (int)b.px[i].v[2] - m.avg.v[2];
            icov_xz += bl * (int)b.px[i].v[0] - m.avg.v[0];
            icov_yz += bl * (int)b.px[i].v[1] - m.avg.v[1];
        }
        if (icov_xz < 0) std::swap(mn.v[0], mx.v[0]);
        if (icov_yz < 0) std::swap(mn.v[1], mx.v[1]);
        r.low = scale_to565(mn);
        r.high = scale_to565(mx);
    } else if (endpoint_mode == QTX_EP_PCA) {
        count(OP_EP_PCA);
        count(OP_PCA_ITER, e.p.power_iterations);
        // BC1Encoder.cpp:606-674
        const float mn[3] = {(float)m.min.v[0], (float)m.min.v[1], (float)m.min.v[2]};
        const float mx[3] = {(float)m.max.v[0], (float)m.max.v[1], (float)m.max.v[2]};
        const float av[3] = {(float)m.avg.v[0], (float)m.avg.v[1], (float)m.avg.v[2]};
        float axis[4] = {306, 601, 117, 0};
        float cov[4][4] = {{1, 0, 0, 0}, {0, 1, 0, 0}, {0, 0, 1, 0}, {0, 0, 0, 1}};
        for (int i = 0; i < 16; i++) {
            if (ignore_black && is_black(b.px[i])) continue;
            float d[3];
            for (int c = 0; c < 3; c++) d[c] = (float)b.px[i].v[c] - av[c];
            for (int c1 = 0; c1 < 3; c1++)
                for (int c2 = c1; c2 < 3; c2++) cov[c1][c2] += (d[c1] * d[c2]);
        }
        for (int rr = 0; rr < 4; rr++)
            for (int c = 0; c < 4; c++) cov[rr][c] = cov[rr][c] / 255.0f;
        for (int rr = 0; rr < 3; rr++)  // Matrix4x4.cpp:41-47
            for (int c = rr + 1; c < 4; c++) cov[c][rr] = cov[rr][c];
        float delta[4] = {mx[0] - mn[0], mx[1] - mn[1], mx[2] - mn[2], 0.0f - 0.0f};
        if (cov[0][2] < 0) delta[0] = -delta[0];
        if (cov[1][2] < 0) delta[1] = -delta[1];
        for (unsigned it = 0; it < e.p.power_iterations; it++) {
            float nd[4];
            for (int rr = 0; rr < 4; rr++) {  // Matrix4x4.cpp:33-39 with Vector4::Dot order
                float s = 0;
                for (int c = 0; c < 4; c++) s += delta[c] * cov[rr][c];
                nd[rr] = s;
            }
            std::memcpy(delta, nd, sizeof nd);
        }
        float k = 0;  // Vector4::MaxAbs(3)
        for (int c = 0; c < 3; c++) {
            const float a = std::fabs(delta[c]);
            if (a > k) k = a;
        }
        if (k >= 2) {
            const float s = 2048.0f / k;
            for (int c = 0; c < 4; c++) axis[c] = delta[c] * s;
        }
        for (int c = 0; c < 4; c++) axis[c] = axis[c] * 16.0f;
        float min_dot = INFINITY, max_dot = -INFINITY;
        int min_index = 0, max_index = 0;
        for (int i = 0; i < 16; i++) {
            if (ignore_black && is_black(b.px[i])) continue;
            const float cv[4] = {(float)b.px[i].v[0], (float)b.px[i].v[1], (float)b.px[i].v[2], 0.0f};
            float dot = 0;
            for (int c = 0; c < 4; c++) dot += cv[c] * axis[c];
            if (dot > max_dot) { max_dot = dot; max_index = i; }
            if (dot < min_dot) { min_dot = dot; min_index = i; }
        }
        r.low = scale_to565(b.px[min_index]);
        r.high = scale_to565(b.px[max_index]);
    }
    r.mode = MODE_INCOMPLETE;
}

// FindSelectors<M>: BC1Encoder.cpp:679-779.  Partial errors and partial
// selector updates on early exit are kept exactly as the reference leaves them.
void find_selectors(const Enc &e, int M, Result &r, const Block &b, int error_mode) {
    const int color_count = M & 0x0F;
    C4 pal[4];
    interpolate_bc1(e.p.interpolator, r.low, r.high, color_count == 3, pal);
    int cv[4][3];
    const int order4[4] = {0, 2, 3, 1}, order3[4] = {0, 2, 1, 3};
    for (int j = 0; j < 4; j++) {
        const C4 &c = pal[color_count == 4 ? order4[j] : order3[j]];
        cv[j][0] = c.v[0]; cv[j][1] = c.v[1]; cv[j][2] = c.v[2];
    }
    unsigned total_error = 0;
    if (error_mode == QTX_ERROR_FULL) count(M == QTX_COLOR_THREE ? OP_SEL_FULL3 : OP_SEL_FULL4);
    else if (error_mode == QTX_ERROR_CHECK2) count(OP_SEL_CHECK2);
    else if (error_mode == QTX_ERROR_FASTER) count(OP_SEL_FASTER);
    else count(OP_SEL_NONE);
    if (error_mode == QTX_ERROR_NONE || error_mode == QTX_ERROR_FASTER) {
        int axis[3] = {cv[3][0] - cv[0][0], cv[3][1] - cv[0][1], cv[3][2] - cv[0][2]};
        int dots[4];
        for (int j = 0; j < 4; j++) dots[j] = dot3(axis, cv[j]);
        const int t0 = dots[0] + dots[1], t1 = dots[1] + dots[2], t2 = dots[2] + dots[3];
        for (int c = 0; c < 3; c++) axis[c] *= 2;
        for (int i = 0; i < 16; i++) {
            const int pv[3] = {b.px[i].v[0], b.px[i].v[1], b.px[i].v[2]};
            const int dot = dot3(axis, pv);
            const int level = (dot <= t0) + (dot < t1) + (dot < t2);
            const int s = 3 - level;
            if (error_mode == QTX_ERROR_FASTER) {
                const int d[3] = {pv[0] - cv[s][0], pv[1] - cv[s][1], pv[2] - cv[s][2]};
                total_error += (unsigned)dot3(d, d);
                if (i % 4 != 0 && total_error >= r.error) break;
            }
            r.sel[i] = (uint8_t)s;
        }
    } else if (error_mode == QTX_ERROR_CHECK2) {
        const int axis[3] = {cv[3][0] - cv[0][0], cv[3][1] - cv[0][1], cv[3][2] - cv[0][2]};
        const float f = 4.0f / ((float)(unsigned)dot3(axis, axis) + .00000125f);
        for (int i = 0; i < 16; i++) {
            const int pv[3] = {b.px[i].v[0], b.px[i].v[1], b.px[i].v[2]};
            const int diff[3] = {pv[0] - cv[0][0], pv[1] - cv[0][1], pv[2] - cv[0][2]};
            const float sel_f = (float)dot3(diff, axis) * f + 0.5f;
            const int s = clampi((int)sel_f, 1, 3);
            const int d0[3] = {cv[s - 1][0] - pv[0], cv[s - 1][1] - pv[1], cv[s - 1][2] - pv[2]};
            const int d1[3] = {cv[s][0] - pv[0], cv[s][1] - pv[1], cv[s][2] - pv[2]};
            const unsigned err0 = (unsigned)dot3(d0, d0), err1 = (unsigned)dot3(d1, d1);
            int best_sel = s;
            unsigned best_err = err1;
            if (err0 == err1) {
                if (best_sel == 1) best_sel = 0;
            } else if (err0 < best_err) {
                best_sel = s - 1;
                best_err = err0;
            }
            total_error += best_err;
            if (total_error >= r.error) break;
            r.sel[i] = (uint8_t)best_sel;
        }
    } else if (error_mode == QTX_ERROR_FULL) {
        const int max_sel = (M == QTX_COLOR_THREE) ? 3 : 4;
        for (int i = 0; i < 16; i++) {
            unsigned best_error = UINT_MAX;
            int best_sel = 0;
            const int pv[3] = {b.px[i].v[0], b.px[i].v[1], b.px[i].v[2]};
            for (int j = 0; j < max_sel; j++) {
                const int d[3] = {cv[j][0] - pv[0], cv[j][1] - pv[1], cv[j][2] - pv[2]};
                const unsigned err = (unsigned)dot3(d, d);
                if (err < best_error || ((err == best_error) && (j == 3))) {
                    best_error = err;
                    best_sel = j;
                }
            }
            total_error += best_error;
            if (total_error >= r.error) break;
            r.sel[i] = (uint8_t)best_sel;
        }
    }
    r.error = total_error;
    r.mode = M;
}

// RefineEndpointsLS(result, pixels, metrics): BC1Encoder.cpp:781-825.
bool refine_endpoints_ls_px(const Enc &e, int M, Result &r, const Block &b, const Metrics &m) {
    count(OP_LS_PX);
    const int color_count = M & 0x0F;
    const Orders &o = color_count == 4 ? g_o4 : g_o3;
    const int denominator = color_count - 1;
    float q00[4] = {0, 0, 0, 0};
    float mat[4] = {0, 0, 0, 0};
    for (int i = 0; i < 16; i++) {
        const C4 &c = b.px[i];
        const int s = r.sel[i];
        // (bool)(M & ThreeColorBlack) and (bool)(M & ThreeColor), as the enum bits
        // test them; the only caller passes FourColor, so neither skip fires.
        if ((M & QTX_COLOR_THREE_BLACK) && is_black(c)) continue;
        if ((M & QTX_COLOR_THREE) && s == 3) continue;
        for (int k = 0; k < 3; k++) q00[k] += (float)(c.v[k] * s);
        for (int k = 0; k < 4; k++) mat[k] += o.weights[s][k];
    }
    const float det = (mat[0] * mat[3]) - (mat[1] * mat[2]);
    if (std::fabs(det) < 1e-8f) {
        r.mode = MODE_INCOMPLETE;
        return false;
    }
    std::swap(mat[0], mat[3]);
    mat[0] = mat[0] * 1.0f; mat[1] = mat[1] * -1.0f; mat[2] = mat[2] * -1.0f; mat[3] = mat[3] * 1.0f;
    const float s = ((float)denominator / 255.0f) / det;
    for (int k = 0; k < 4; k++) mat[k] = mat[k] * s;
    float q10[3], lo[3], hi[3];
    for (int k = 0; k < 3; k++) q10[k] = (float)(m.sums[k] * denominator) - q00[k];
    for (int k = 0; k < 3; k++) {
        lo[k] = (mat[0] * q00[k]) + (mat[1] * q10[k]);
        hi[k] = (mat[2] * q00[k]) + (mat[3] * q10[k]);
    }
    r.mode = M;
    r.low = precise_round565(lo);
    r.high = precise_round565(hi);
    return true;
}

// RefineEndpointsLS(result, sums, matrix, hash): BC1Encoder.cpp:827-850.
void refine_endpoints_ls_sums(int M, Result &r, const float sums[17][3], const float mat[4], const Orders &o,
                              uint16_t hash) {
    count(OP_LS_SUMS);
    const int color_count = M & 0x0F;
    const int denominator = color_count - 1;
    float q10[3] = {0, 0, 0};
    unsigned level = 0;
    const uint8_t *h = o.hist + hash * o.n;
    for (int i = 0; i < color_count - 1; i++) {
        level += h[i];
        for (int k = 0; k < 3; k++) q10[k] += sums[level][k];
    }
    float q00[3], lo[3], hi[3];
    for (int k = 0; k < 3; k++) q00[k] = (sums[16][k] * (float)denominator) - q10[k];
    for (int k = 0; k < 3; k++) {
        lo[k] = (mat[0] * q00[k]) + (mat[1] * q10[k]);
        hi[k] = (mat[2] * q00[k]) + (mat[3] * q10[k]);
    }
    r.mode = M;
    r.low = precise_round565(lo);
    r.high = precise_round565(hi);
}

// RefineBlockLS<M>: BC1Encoder.cpp:852-875 (always solves with 4-colour weights, :860)
void refine_block_ls(const Enc &e, int M, Result &r, const Block &b, const Metrics &m, int error_mode, unsigned passes) {
    for (unsigned pass = 0; pass < passes; pass++) {
        Result t = r;
        const bool multicolor = refine_endpoints_ls_px(e, QTX_COLOR_FOUR, t, b, m);
        if (!multicolor) endpoints_single_color_px(e, t, b, m.avg, M != QTX_COLOR_FOUR);
        else find_selectors(e, M, t, b, error_mode);
        if (eq(t.low, r.low) && eq(t.high, r.high)) break;
        if (error_mode == QTX_ERROR_NONE || t.error < r.error) r = t;
        else return;
    }
}

// RefineBlockCF<M>: BC1Encoder.cpp:877-928
void refine_block_cf(const Enc &e, int M, Result &r, const Block &b, const Metrics &m, int error_mode,
                     unsigned orderings) {
    const int color_count = M & 0x0F;
    const Orders &o = color_count == 4 ? g_o4 : g_o3;
    count(OP_CF_SETUP);
    const Result orig = r;
    uint8_t h[4] = {0, 0, 0, 0};
    for (int i = 0; i < 16; i++) h[orig.sel[i]]++;
    const uint16_t start_hash = get_hash(o, h);
    const C4 hi8 = scale_from565(orig.high), lo8 = scale_from565(orig.low);
    const float axis[4] = {(float)((int)hi8.v[0] - lo8.v[0]), (float)((int)hi8.v[1] - lo8.v[1]),
                           (float)((int)hi8.v[2] - lo8.v[2]), (float)((int)hi8.v[3] - lo8.v[3])};
    float cvec[16][4];
    uint32_t dots[16];
    for (int i = 0; i < 16; i++) {
        for (int c = 0; c < 3; c++) cvec[i][c] = (float)b.px[i].v[c];
        cvec[i][3] = 0;
        float d = 0;
        for (int c = 0; c < 4; c++) d += cvec[i][c] * axis[c];
        const int dot = 0x1000000 + (int)d;
        dots[i] = (uint32_t)(dot << 4) | (uint32_t)i;
    }
    std::sort(dots, dots + 16);
    float sums[17][3];
    sums[0][0] = sums[0][1] = sums[0][2] = 0;
    for (int i = 0; i < 16; i++) {
        const unsigned p = dots[i] & 0xF;
        for (int c = 0; c < 3; c++) sums[i + 1][c] = sums[i][c] + cvec[p][c];
    }
    const unsigned q_total = e.p.exhaustive ? (unsigned)o.count : orderings;
    for (unsigned q = 0; q < q_total; q++) {
        const uint16_t trial_hash = e.p.exhaustive ? (uint16_t)q : o.best[start_hash * o.best_count + q];
        Result t = orig;
        if (is_single_color_hash(o, trial_hash)) {
            endpoints_single_color_px(e, t, b, m.avg, color_count == 3);
        } else {
            refine_endpoints_ls_sums(M, t, sums, o.factors[trial_hash], o, trial_hash);
            find_selectors(e, M, t, b, error_mode);
        }
        if (t.error < r.error) r = t;
        if (t.error == 0) break;
    }
}

// EndpointSearch: BC1Encoder.cpp:930-996
const int kVoxels[16][4] = {{1, 0, 0, 3},  {0, 1, 0, 4},   {0, 0, 1, 5},   {-1, 0, 0, 0}, {0, -1, 0, 1}, {0, 0, -1, 2},
                            {1, 1, 0, 9},  {1, 0, 1, 10},  {0, 1, 1, 11},  {-1, -1, 0, 6}, {-1, 0, -1, 7}, {0, -1, -1, 8},
                            {-1, 1, 0, 13}, {1, -1, 0, 12}, {0, -1, 1, 15}, {0, 1, -1, 14}};

void endpoint_search(const Enc &e, Result &r, const Block &b) {
    unsigned prev_improvement_index = 0;
    int forbidden_direction = -1;
    for (unsigned i = 0; i < e.p.search_rounds; i++) {
        const unsigned vi = i & 15;
        if ((int)(i & 31) == forbidden_direction) continue;
        const int *d = kVoxels[vi];
        count(OP_SEARCH_ROUND);
        Result t = r;
        C4 &ep = (i & 16) ? t.low : t.high;
        ep.v[0] = (uint8_t)clampi(ep.v[0] + d[0], 0, 31);
        ep.v[1] = (uint8_t)clampi(ep.v[1] + d[1], 0, 63);
        ep.v[2] = (uint8_t)clampi(ep.v[2] + d[2], 0, 31);
        switch (r.mode) {
            case QTX_COLOR_THREE: find_selectors(e, QTX_COLOR_THREE, t, b, QTX_ERROR_FULL); break;
            case QTX_COLOR_THREE_BLACK: find_selectors(e, QTX_COLOR_THREE_BLACK, t, b, QTX_ERROR_FULL); break;
            default: find_selectors(e, QTX_COLOR_FOUR, t, b, e.p.error_mode); break;
        }
        if (t.error < r.error) {
            r = t;
            forbidden_direction = d[3] | (int)(i & 16);
            prev_improvement_index = i;
        }
        if (i - prev_improvement_index > 32) break;
    }
}

void put_block(uint8_t out[8], uint16_t ep0, uint16_t ep1, const uint8_t rows[4]) {
    out[0] = (uint8_t)(ep0 & 0xFF); out[1] = (uint8_t)(ep0 >> 8);
    out[2] = (uint8_t)(ep1 & 0xFF); out[3] = (uint8_t)(ep1 >> 8);
    for (int y = 0; y < 4; y++) out[4 + y] = rows[y];
}

// WriteBlockSolid: BC1Encoder.cpp:345-393
void write_block_solid(const Enc &e, const C4 &color, uint8_t out[8]) {
    count(OP_SOLID);
    uint8_t mask = 0xAA;
    uint16_t min16, max16;
    if ((color.v[0] | color.v[1] | color.v[2]) == 0) {
        min16 = 0; max16 = 1; mask = 0x55;
    } else {
        Result res;
        endpoints_single_color(e, res, color, false);
        if (e.p.color_mode & QTX_COLOR_THREE) {
            Result r3;
            endpoints_single_color(e, r3, color, true);
            if (r3.error < res.error) res = r3;
        }
        min16 = pack565u(res.low);
        max16 = pack565u(res.high);
        if (res.mode == QTX_COLOR_FOUR) {
            if (min16 == max16) {
                if (min16 > 0) { min16--; mask = 0; }
                else { max16 = 1; min16 = 0; mask = 0x55; }
            } else if (max16 < min16) {
                std::swap(min16, max16);
                mask = 0xFF;
            }
        } else if (max16 > min16) {
            std::swap(min16, max16);
        }
    }
    const uint8_t rows[4] = {mask, mask, mask, mask};
    put_block(out, max16, min16, rows);
}

// WriteBlock: BC1Encoder.cpp:395-441
void write_block(const Result &r, uint8_t out[8]) {
    count(OP_WRITE);
    uint16_t ep1 = pack565u(r.low), ep0 = pack565u(r.high);
    uint8_t lut[4];
    if (r.mode & QTX_COLOR_FOUR) {
        const uint8_t l0[4] = {1, 3, 2, 0};
        std::memcpy(lut, l0, 4);
        if (ep1 > ep0) {
            std::swap(ep1, ep0);
            const uint8_t l1[4] = {0, 2, 3, 1};
            std::memcpy(lut, l1, 4);
        } else if (ep1 == ep0) {
            if (ep1 > 0) {
                ep1--;
                std::memset(lut, 0, 4);
            } else {
                ep0 = 1; ep1 = 0;
                std::memset(lut, 1, 4);
            }
        }
    } else {
        const uint8_t l0[4] = {1, 2, 0, 3};
        std::memcpy(lut, l0, 4);
        if (ep1 < ep0) {
            std::swap(ep1, ep0);
            const uint8_t l1[4] = {0, 2, 1, 3};
            std::memcpy(lut, l1, 4);
        }
    }
    uint8_t rows[4] = {0, 0, 0, 0};
    for (int i = 0; i < 16; i++) rows[i / 4] |= (uint8_t)(lut[r.sel[i]] << (2 * (i % 4)));
    put_block(out, ep0, ep1, rows);
}

// EncodeBlock: BC1Encoder.cpp:261-342
void encode_bc1_block(const Enc &e, const Block &b, uint8_t out[8]) {
    if (is_single_color(b)) {
        write_block_solid(e, b.px[0], out);
        return;
    }
    const Metrics m = get_metrics(b, false);
    const bool use_likely_orderings = e.p.exhaustive || e.p.orderings3 > 0 || e.p.orderings4 > 0;
    bool needs_block_error = use_likely_orderings;
    needs_block_error |= (e.p.color_mode == QTX_COLOR_THREE);
    needs_block_error |= (e.p.color_mode == QTX_COLOR_THREE_BLACK) && m.has_black;
    needs_block_error |= (e.p.error_mode != QTX_ERROR_NONE);
    needs_block_error |= (e.p.search_rounds > 0);
    const int error_mode = needs_block_error ? e.p.error_mode : QTX_ERROR_NONE;
    const unsigned total_ls_passes = e.p.two_ls_passes ? 2 : 1;
    const unsigned total_cf_passes = e.p.two_cf_passes ? 2 : 1;
    const unsigned total_ep_passes = (needs_block_error && e.p.two_ep_passes) ? 2 : 1;

    Result orig, result;
    for (unsigned round = 0; round < total_ep_passes; round++) {
        const int endpoint_mode = (round == 1) ? QTX_EP_BOUNDING_BOX : e.p.endpoint_mode;
        Result trial_orig;
        find_endpoints(e, trial_orig, b, m, endpoint_mode, false);
        Result trial = trial_orig;
        find_selectors(e, QTX_COLOR_FOUR, trial, b, error_mode);
        refine_block_ls(e, QTX_COLOR_FOUR, trial, b, m, error_mode, total_ls_passes);
        if (!needs_block_error || trial.error < result.error) {
            result = trial;
            orig = trial_orig;
        }
    }
    if (result.error > 0 && use_likely_orderings) {
        for (unsigned it = 0; it < total_cf_passes; it++)
            refine_block_cf(e, QTX_COLOR_FOUR, result, b, m, e.p.error_mode, e.p.orderings4);
    }
    if (result.error > 0 && (e.p.color_mode & QTX_COLOR_THREE)) {
        Result trial = orig;
        find_selectors(e, QTX_COLOR_THREE, trial, b, QTX_ERROR_FULL);
        refine_block_ls(e, QTX_COLOR_THREE, trial, b, m, QTX_ERROR_FULL, total_ls_passes);
        if (trial.error > 0 && use_likely_orderings) {
            for (unsigned it = 0; it < total_cf_passes; it++)
                refine_block_cf(e, QTX_COLOR_THREE, trial, b, m, QTX_ERROR_FULL, e.p.orderings3);
        }
        if (trial.error < result.error) result = trial;
    }
    if (result.error > 0 && e.p.color_mode == QTX_COLOR_THREE_BLACK && m.has_black && !is_black(m.max)) {
        Result trial;
        const Metrics mnb = get_metrics(b, true);
        find_endpoints(e, trial, b, mnb, QTX_EP_PCA, true);
        find_selectors(e, QTX_COLOR_THREE_BLACK, trial, b, QTX_ERROR_FULL);
        refine_block_ls(e, QTX_COLOR_THREE_BLACK, trial, b, mnb, QTX_ERROR_FULL, total_ls_passes);
        if (trial.error < result.error) result = trial;
    }
    if (result.error > 0 && e.p.search_rounds > 0) endpoint_search(e, result, b);
    write_block(result, out);
}

// ----------------------------------------------------------------------------
// BC4: quicktex/s3tc/bc4/BC4Encoder.cpp:30-71, BC4Block.cpp:34-64

void put_bc4(uint8_t out[8], int a0, int a1, const uint8_t sel[16]) {
    out[0] = (uint8_t)a0;
    out[1] = (uint8_t)a1;
    uint64_t bits = 0;
    for (int i = 0; i < 16; i++) bits |= (uint64_t)(sel[i] & 7) << (3 * i);
    for (int k = 0; k < 6; k++) out[2 + k] = (uint8_t)(bits >> (8 * k));
}

void encode_bc4_block(const Block &b, int ch, uint8_t out[8]) {
    count(OP_BC4);
    int mn = 255, mx = 0;
    for (int i = 0; i < 16; i++) {
        mn = std::min(mn, (int)b.px[i].v[ch]);
        mx = std::max(mx, (int)b.px[i].v[ch]);
    }
    uint8_t sel[16] = {0};
    if (mx == mn) {
        put_bc4(out, mn, mn, sel);
        return;
    }
    static const uint8_t levels[8] = {1, 7, 6, 5, 4, 3, 2, 0};
    const int bias = 4 - mn * 14;
    const int delta = mx - mn;
    int thr[7];
    for (int i = 0; i < 7; i++) thr[i] = delta * (1 + 2 * i) - bias;
    for (int i = 0; i < 16; i++) {
        const int value = (int)b.px[i].v[ch] * 14;
        int level = 0;
        for (int c = 0; c < 7; c++) level += value >= thr[c];
        sel[i] = levels[level];
    }
    put_bc4(out, mx, mn, sel);
}

void bc4_values(int a0, int a1, uint8_t v[8]) {
    v[0] = (uint8_t)a0;
    v[1] = (uint8_t)a1;
    if (a0 <= a1) {  // BC4Block.cpp:44-53
        v[2] = (uint8_t)((a0 * 4 + a1) / 5);
        v[3] = (uint8_t)((a0 * 3 + a1 * 2) / 5);
        v[4] = (uint8_t)((a0 * 2 + a1 * 3) / 5);
        v[5] = (uint8_t)((a0 + a1 * 4) / 5);
        v[6] = 0;
        v[7] = 0xFF;
    } else {  // BC4Block.cpp:55-64
        for (int k = 1; k <= 6; k++) v[1 + k] = (uint8_t)((a0 * (7 - k) + a1 * k) / 7);
    }
}

// BC4Decoder::DecodeInto: BC4Decoder.cpp:29-43
void decode_bc4_into(const uint8_t in[8], int ch, Block &dst) {
    uint8_t v[8];
    bc4_values(in[0], in[1], v);
    uint64_t bits = 0;
    for (int k = 0; k < 6; k++) bits |= (uint64_t)in[2 + k] << (8 * k);
    for (int i = 0; i < 16; i++) dst.px[i].v[ch] = v[(bits >> (3 * i)) & 7];
}

// BC1Decoder::DecodeBlock: BC1Decoder.cpp:31-52
void decode_bc1_into(const uint8_t in[8], int interp, bool use_3color, bool write_alpha, Block &dst) {
    const uint16_t c0 = (uint16_t)(in[0] | (in[1] << 8));
    const uint16_t c1 = (uint16_t)(in[2] | (in[3] << 8));
    const bool three = use_3color && (c1 >= c0);  // Interpolator.cpp:52-55
    const C4 lo = mk((c0 >> 11) & 31, (c0 >> 5) & 63, c0 & 31);
    const C4 hi = mk((c1 >> 11) & 31, (c1 >> 5) & 63, c1 & 31);
    C4 pal[4];
    interpolate_bc1(interp, lo, hi, three, pal);
    for (int i = 0; i < 16; i++) {
        const int s = (in[4 + i / 4] >> (2 * (i % 4))) & 3;
        C4 c = pal[s];
        if (!write_alpha) c.v[3] = dst.px[i].v[3];
        dst.px[i] = c;
    }
}

// ----------------------------------------------------------------------------
// Texture drivers: quicktex/Encoder.h:36-56, Decoder.h:36-54, Texture.h:89-134

Block gather(const uint8_t *rgba, int w, int h, int bx, int by) {
    Block b;
    for (int y = 0; y < 4; y++)
        for (int x = 0; x < 4; x++) {
            const int px = (bx * 4 + x) % w, py = (by * 4 + y) % h;
            std::memcpy(b.px[x + 4 * y].v, rgba + 4 * ((size_t)py * w + px), 4);
        }
    return b;
}

void scatter(uint8_t *rgba, int w, int h, int bx, int by, const Block &b) {
    for (int x = 0; x < 4; x++)
        for (int y = 0; y < 4; y++) {
            const int px = (bx * 4 + x) % w, py = (by * 4 + y) % h;
            std::memcpy(rgba + 4 * ((size_t)py * w + px), b.px[x + 4 * y].v, 4);
        }
}

Block blank_block() {
    Block b;
    for (int i = 0; i < 16; i++) b.px[i] = mk(0, 0, 0, 255);
    return b;
}

thread_local char g_err[256];
int g_threads = 0;  // 0: mirror the reference's OpenMP policy

int fail(const char *msg) {
    std::snprintf(g_err, sizeof g_err, "%s", msg);
    return 1;
}

int check_dims(int w, int h) {
    if (w <= 0) return fail("Texture width must be greater than 0");
    if (h <= 0) return fail("Texture height must be greater than 0");
    return 0;
}

template <class F> void for_blocks(int w, int h, bool parallel, F &&f) {
    const int bw = (w + 3) / 4, bh = (h + 3) / 4;
    const bool par = parallel && (bw * bh >= 16);
    if (par) {
        if (g_threads > 0) {
#pragma omp parallel for num_threads(g_threads)
            for (int y = 0; y < bh; y++)
                for (int x = 0; x < bw; x++) f(x, y);
        } else {
#pragma omp parallel for
            for (int y = 0; y < bh; y++)
                for (int x = 0; x < bw; x++) f(x, y);
        }
    } else {
        for (int y = 0; y < bh; y++)
            for (int x = 0; x < bw; x++) f(x, y);
    }
}

bool all_parallel() { return g_threads > 0; }

}  // namespace

extern "C" {

const char *qtxo_last_error(void) { return g_err; }

// Threads for every format: 0 mirrors the reference (OpenMP over block rows
// for BC1 with >= 16 blocks, Encoder.h:46 / BC1Encoder.h:135; serial otherwise),
// n > 0 uses n threads for every encoder.
void qtxo_set_threads(int n) { g_threads = n; }

int qtxo_bc1_params_init(qtx_bc1_params *p, uint32_t level, int32_t color_mode, int32_t interp) {
    // BC1Encoder.cpp:47-50 and SetLevel :74-249
    if (color_mode != QTX_COLOR_FOUR && color_mode != QTX_COLOR_THREE && color_mode != QTX_COLOR_THREE_BLACK)
        return fail("Encoder color mode must be FourColor, ThreeColor, or ThreeColorBlack");
    if (interp < 0 || interp > 3) return fail("Invalid interpolator type");
    if (level > 19) return fail("Level out of range, bust be between 0 and 18 inclusive");
    p->interpolator = interp;
    p->color_mode = color_mode;
    p->two_ls_passes = p->two_ep_passes = p->two_cf_passes = p->exhaustive = 0;
    p->power_iterations = 4;
    p->error_mode = QTX_ERROR_CHECK2;
    p->endpoint_mode = QTX_EP_PCA;
    p->search_rounds = 0;
    unsigned o3 = 0, o4 = 0;
    switch (level) {
        case 0: p->endpoint_mode = QTX_EP_BOUNDING_BOX_INT; break;
        case 1: p->endpoint_mode = QTX_EP_LEAST_SQUARES; break;
        case 2: break;
        case 3: p->two_ls_passes = 1; break;
        case 4: p->two_ls_passes = 1; p->error_mode = QTX_ERROR_FULL; p->power_iterations = 6; break;
        default:
        case 5: p->two_ls_passes = 1; p->error_mode = QTX_ERROR_FASTER; break;
        case 6: p->two_ls_passes = 1; o4 = 1; o3 = 1; p->error_mode = QTX_ERROR_FASTER; break;
        case 7: p->two_ls_passes = 1; p->error_mode = QTX_ERROR_FASTER; o4 = 4; o3 = 1; break;
        case 8: p->two_ls_passes = 1; p->error_mode = QTX_ERROR_FASTER; o4 = 8; o3 = 1; break;
        case 9: p->two_ls_passes = 1; p->error_mode = QTX_ERROR_CHECK2; o4 = 11; o3 = 3; break;
        case 10: p->two_ls_passes = 1; p->error_mode = QTX_ERROR_CHECK2; o4 = 20; o3 = 8; break;
        case 11: p->two_ls_passes = 1; p->error_mode = QTX_ERROR_CHECK2; o4 = 28; o3 = 16; break;
        case 12: p->two_ls_passes = 1; p->error_mode = QTX_ERROR_CHECK2; o4 = 32; o3 = 32; break;
        case 13: p->two_ls_passes = p->two_ep_passes = 1; p->error_mode = QTX_ERROR_FULL; o4 = 32; o3 = 32;
                 p->search_rounds = 20; p->power_iterations = 6; break;
        case 14: p->two_ls_passes = p->two_ep_passes = 1; p->error_mode = QTX_ERROR_FULL; o4 = 32; o3 = 32;
                 p->search_rounds = 32; p->power_iterations = 6; break;
        case 15: p->two_ls_passes = p->two_ep_passes = 1; p->error_mode = QTX_ERROR_FULL; o4 = 56; o3 = 32;
                 p->search_rounds = 32; p->power_iterations = 6; break;
        case 16: p->two_ls_passes = p->two_ep_passes = 1; p->error_mode = QTX_ERROR_FULL; o4 = 80; o3 = 32;
                 p->search_rounds = 256; p->power_iterations = 6; break;
        case 17: p->two_ls_passes = p->two_ep_passes = 1; p->error_mode = QTX_ERROR_FULL; o4 = 128; o3 = 32;
                 p->search_rounds = 256; break;
        case 18: p->two_ls_passes = p->two_ep_passes = p->two_cf_passes = 1; p->error_mode = QTX_ERROR_FULL;
                 o4 = 128; o3 = 32; p->search_rounds = 256; p->power_iterations = 6; break;
        case 19: p->two_ls_passes = p->two_ep_passes = p->two_cf_passes = p->exhaustive = 1;
                 p->error_mode = QTX_ERROR_FULL; o4 = 128; o3 = 32; p->search_rounds = 256; p->power_iterations = 6;
                 break;
    }
    p->orderings4 = (unsigned)clampi((int)o4, 1, 128);
    p->orderings3 = (unsigned)clampi((int)o3, 1, 32);
    return 0;
}

int qtxo_bc1_encode(const qtx_bc1_params *p, const uint8_t *rgba, int32_t w, int32_t h, uint8_t *out) {
    if (check_dims(w, h)) return 1;
    init_tables();
    Enc e{*p};
    const int bw = (w + 3) / 4;
    for_blocks(w, h, true, [&](int x, int y) {
        const Block b = gather(rgba, w, h, x, y);
        encode_bc1_block(e, b, out + 8 * ((size_t)y * bw + x));
    });
    return 0;
}

int qtxo_bc3_encode(const qtx_bc1_params *p, const uint8_t *rgba, int32_t w, int32_t h, uint8_t *out) {
    if (check_dims(w, h)) return 1;
    if (p->color_mode != QTX_COLOR_FOUR) return fail("BC3 colour blocks are always FourColor");
    init_tables();
    Enc e{*p};
    const int bw = (w + 3) / 4;
    for_blocks(w, h, all_parallel(), [&](int x, int y) {
        const Block b = gather(rgba, w, h, x, y);
        uint8_t *o = out + 16 * ((size_t)y * bw + x);
        encode_bc1_block(e, b, o + 8);  // BC3Encoder.cpp:27-32; alpha block first (BC3Block.h:35-36)
        encode_bc4_block(b, 3, o);
    });
    return 0;
}

int qtxo_bc4_encode(uint32_t ch, const uint8_t *rgba, int32_t w, int32_t h, uint8_t *out) {
    if (ch >= 4) return fail("Channel out of range");
    if (check_dims(w, h)) return 1;
    const int bw = (w + 3) / 4;
    for_blocks(w, h, all_parallel(), [&](int x, int y) {
        encode_bc4_block(gather(rgba, w, h, x, y), (int)ch, out + 8 * ((size_t)y * bw + x));
    });
    return 0;
}

int qtxo_bc5_encode(uint32_t c0, uint32_t c1, const uint8_t *rgba, int32_t w, int32_t h, uint8_t *out) {
    if (c0 >= 4 || c1 >= 4) return fail("Channel out of range");
    if (check_dims(w, h)) return 1;
    const int bw = (w + 3) / 4;
    for_blocks(w, h, all_parallel(), [&](int x, int y) {
        const Block b = gather(rgba, w, h, x, y);
        uint8_t *o = out + 16 * ((size_t)y * bw + x);
        encode_bc4_block(b, (int)c0, o);
        encode_bc4_block(b, (int)c1, o + 8);
    });
    return 0;
}

// Decoders are serial in block order (Decoder.h:45-51): later blocks overwrite
// wrapped texels of earlier ones.
int qtxo_bc1_decode(int32_t interp, int32_t write_alpha, const uint8_t *blocks, int32_t w, int32_t h, uint8_t *out) {
    if (check_dims(w, h)) return 1;
    init_tables();
    const int bw = (w + 3) / 4, bh = (h + 3) / 4;
    for (int y = 0; y < bh; y++)
        for (int x = 0; x < bw; x++) {
            Block b = blank_block();
            decode_bc1_into(blocks + 8 * ((size_t)y * bw + x), interp, true, write_alpha != 0, b);
            scatter(out, w, h, x, y, b);
        }
    return 0;
}

int qtxo_bc3_decode(int32_t interp, const uint8_t *blocks, int32_t w, int32_t h, uint8_t *out) {
    if (check_dims(w, h)) return 1;
    init_tables();
    const int bw = (w + 3) / 4, bh = (h + 3) / 4;
    for (int y = 0; y < bh; y++)
        for (int x = 0; x < bw; x++) {
            const uint8_t *in = blocks + 16 * ((size_t)y * bw + x);
            Block b = blank_block();
            decode_bc1_into(in + 8, interp, false, false, b);  // BC3Decoder.cpp:26-32
            decode_bc4_into(in, 3, b);
            scatter(out, w, h, x, y, b);
        }
    return 0;
}

int qtxo_bc4_decode(uint32_t ch, const uint8_t *blocks, int32_t w, int32_t h, uint8_t *out) {
    if (ch >= 4) return fail("Channel out of range");
    if (check_dims(w, h)) return 1;
    const int bw = (w + 3) / 4, bh = (h + 3) / 4;
    for (int y = 0; y < bh; y++)
        for (int x = 0; x < bw; x++) {
            Block b = blank_block();
            decode_bc4_into(blocks + 8 * ((size_t)y * bw + x), (int)ch, b);
            scatter(out, w, h, x, y, b);
        }
    return 0;
}

int qtxo_bc5_decode(uint32_t c0, uint32_t c1, const uint8_t *blocks, int32_t w, int32_t h, uint8_t *out) {
    if (c0 >= 4 || c1 >= 4) return fail("Channel out of range");
    if (check_dims(w, h)) return 1;
    const int bw = (w + 3) / 4, bh = (h + 3) / 4;
    for (int y = 0; y < bh; y++)
        for (int x = 0; x < bw; x++) {
            const uint8_t *in = blocks + 16 * ((size_t)y * bw + x);
            Block b = blank_block();
            decode_bc4_into(in, (int)c0, b);
            decode_bc4_into(in + 8, (int)c1, b);
            scatter(out, w, h, x, y, b);
        }
    return 0;
}

// Table access for the table pins in tests/.
int qtxo_match_table(int32_t interp, int32_t bits, uint8_t *out /* 256 x {high, low, error} */) {
    if (interp < 0 || interp > 3 || (bits != 5 && bits != 6)) return fail("bad table");
    init_tables();
    std::memcpy(out, g_match[interp][bits == 6], 256 * 3);
    return 0;
}

int qtxo_order_factors(int32_t n, float *out /* count x 4 */) {
    if (n != 3 && n != 4) return fail("bad table");
    init_tables();
    const Orders &o = n == 4 ? g_o4 : g_o3;
    std::memcpy(out, o.factors, sizeof(float) * 4 * o.count);
    return 0;
}

// Algorithmic op accounting of the calling thread (see OpKind).
int qtxo_op_kinds(void) { return OP_KINDS; }
void qtxo_op_reset(void) { std::memset(g_ops, 0, sizeof g_ops); }
void qtxo_op_counts(uint64_t *out) { std::memcpy(out, g_ops, sizeof g_ops); }
void qtxo_op_weights(uint32_t *out) { std::memcpy(out, kOpWeight, sizeof kOpWeight); }

int qtxo_midpoints(float *out5 /* 32 */, float *out6 /* 64 */) {
    init_tables();
    std::memcpy(out5, g_mid5, sizeof g_mid5);
    std::memcpy(out6, g_mid6, sizeof g_mid6);
    return 0;
}

}  // extern "C"

// ----------------------------------------------------------------------------
// Mip levels of quicktex.dds.encode (quicktex/dds.py:280): every level is
// resampled from the base image by image_utils.resize_no_premultiply
// (quicktex/image_utils.py:42-52), i.e. Pillow's Image.resize with
// Image.BILINEAR on the RGB bands and, separately, on the alpha band.  Pillow
// (a dependency of the reference, unpinned in pyproject.toml:30; 12.2.0 in this
// image) implements it in libImaging/Resample.c; restated here from its
// published algorithm:
//   * per axis, coefficients of the triangle filter (support 1) scaled by the
//     downscale factor, centred at in0 + (i + 0.5) * scale, each window
//     [xmin, xmin + n) clipped to the image and normalised to sum 1, in double;
//   * the weights as fixed point with 22 fraction bits, rounded half away
//     from zero;
//   * a horizontal pass into an 8-bit image, then a vertical pass, each output
//     = clamp((2^21 + sum(v * k)) >> 22, 0, 255); a pass is skipped when that
//     axis keeps its size.
// Bands are resampled independently with the same coefficients, so resampling
// all four channels of RGBA8 equals resize_no_premultiply on RGBA and
// Image.resize on RGBX.  Pinned against Pillow itself in tests/test_oracle.py.
namespace {

constexpr int kResampleBits = 22;

struct Coeffs {
    int ksize = 0;
    std::vector<int32_t> k;      // out x ksize
    std::vector<int32_t> lo, n;  // window start and length per output
};

Coeffs resample_coeffs(int in_size, int out_size) {
    Coeffs c;
    const double scale = (double)in_size / (double)out_size;
    const double filterscale = scale < 1.0 ? 1.0 : scale;
    const double support = 1.0 * filterscale;
    c.ksize = (int)std::ceil(support) * 2 + 1;
    c.k.assign((size_t)out_size * c.ksize, 0);
    c.lo.resize(out_size);
    c.n.resize(out_size);
    std::vector<double> w(c.ksize);
    for (int xx = 0; xx < out_size; xx++) {
        const double center = (xx + 0.5) * scale;
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
            w[x] = t < 1.0 ? 1.0 - t : 0.0;
            ww += w[x];
        }
        for (int x = 0; x < xmax; x++) {
            const double v = ww != 0.0 ? w[x] / ww : w[x];
            c.k[(size_t)xx * c.ksize + x] =
                v < 0 ? (int32_t)(-0.5 + v * (1 << kResampleBits)) : (int32_t)(0.5 + v * (1 << kResampleBits));
        }
        c.lo[xx] = xmin;
        c.n[xx] = xmax;
    }
    return c;
}

inline uint8_t clip8(int32_t v) {
    v >>= kResampleBits;
    return (uint8_t)(v < 0 ? 0 : v > 255 ? 255 : v);
}

}  // namespace

extern "C" int qtxo_resize(const uint8_t *rgba, int32_t w, int32_t h, uint8_t *out, int32_t ow, int32_t oh) {
    if (w <= 0 || h <= 0 || ow <= 0 || oh <= 0) return fail("bad size");
    std::vector<uint8_t> tmp;
    const uint8_t *src = rgba;
    int sw = w;
    if (ow != w) {  // horizontal pass
        const Coeffs c = resample_coeffs(w, ow);
        tmp.resize((size_t)ow * h * 4);
        for (int y = 0; y < h; y++)
            for (int x = 0; x < ow; x++)
                for (int ch = 0; ch < 4; ch++) {
                    int32_t ss = 1 << (kResampleBits - 1);
                    for (int t = 0; t < c.n[x]; t++)
                        ss += rgba[((size_t)y * w + c.lo[x] + t) * 4 + ch] * c.k[(size_t)x * c.ksize + t];
                    tmp[((size_t)y * ow + x) * 4 + ch] = clip8(ss);
                }
        src = tmp.data();
        sw = ow;
    }
    if (oh != h) {  // vertical pass
        const Coeffs c = resample_coeffs(h, oh);
        for (int y = 0; y < oh; y++)
            for (int x = 0; x < sw; x++)
                for (int ch = 0; ch < 4; ch++) {
                    int32_t ss = 1 << (kResampleBits - 1);
                    for (int t = 0; t < c.n[y]; t++)
                        ss += src[((size_t)(c.lo[y] + t) * sw + x) * 4 + ch] * c.k[(size_t)y * c.ksize + t];
                    out[((size_t)y * sw + x) * 4 + ch] = clip8(ss);
                }
    } else {
        std::memcpy(out, src, (size_t)sw * h * 4);
    }
    return 0;
}
