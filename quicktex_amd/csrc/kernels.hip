// kernels.hip -- gfx950 kernels of the S3TC path and their launchers.
//
// Layouts (quicktex/Texture.h:73-87, 158-162): raw textures are row-major RGBA8
// with no row padding; block textures are row-major blocks of 8 (BC1/BC4) or
// 16 (BC3/BC5) bytes.
//
// Edge rules (quicktex/Texture.h:89-134):
//   encode gathers texel (tx, ty) of block (bx, by) from ((4bx+tx) mod w, (4by+ty) mod h);
//   decode writes blocks in row-major order, texels x-outer / y-inner, so a
//   texture pixel keeps the value of the last block texel that maps onto it.
//   Per axis that is the last texel of the last block if the last block wraps
//   onto the pixel, else the pixel's own position; decode writes only those
//   texels, so no two lanes write one pixel.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "bc1_device.h"
#include "bc1_lane.h"
#include "kernels.h"

namespace qtx {
namespace dev {

// ---------------------------------------------------------------------------
// Gather of one 4x4 block by one lane per texel (lane i of the group loads
// texel i) then shared across the group.
__device__ __forceinline__ uint32_t load_texel(const uint8_t *__restrict__ rgba, int w, int h, int bx, int by, int t) {
    int x = bx * 4 + (t & 3), y = by * 4 + (t >> 2);
    if (x >= w) x %= w;
    if (y >= h) y %= h;
    return *reinterpret_cast<const uint32_t *>(rgba + 4 * ((size_t)y * (size_t)w + (size_t)x));
}

// High 32 bits of the product of two 24-bit values (v_mul_hi_u32_u24, a full-rate
// instruction; the 32-bit v_mul_hi_u32 is not).
__device__ __forceinline__ uint32_t mul_hi_u24(uint32_t a, uint32_t b) {
    uint32_t r;
    asm("v_mul_hi_u32_u24 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
    return r;
}

// BC4Encoder::EncodeBlock (quicktex/s3tc/bc4/BC4Encoder.cpp:30-71); v[i] in [0,255].
__device__ __forceinline__ uint2 bc4_encode_values(const uint32_t v[16]) {
    uint32_t mn = 255, mx = 0;
#pragma unroll
    for (int i = 0; i < 16; i++) {
        mn = min(mn, v[i]);
        mx = max(mx, v[i]);
    }
    if (mx == mn) return make_uint2(mn | (mn << 8), 0u);
    const uint32_t delta = mx - mn, d = 2u * delta;
    // The reference's level is the number of c in 0..6 with 14 v >= delta (1 + 2c) -
    // (4 - 14 mn), that is y >= d (c + 1) for y = 14 (v - mn) + 4 + delta and
    // d = 2 delta: so min(floor(y / d), 7).  This takes q = floor(y / d) + 1 =
    // floor(y3 / d), y3 = y + d in [7, 4339], as (y3 m) >> 22 with m =
    // ceil(2^22 / d): y3 m / 2^22 exceeds y3 / d by less than 4339 / 2^22 = 0.00103,
    // below the gap 1 / 510 between a quotient that is no integer and the next one
    // -- checked for every delta and value on the CPU.  In 24-bit arithmetic:
    // y3 2^10 < 2^24 and m <= 2^21, so q is the high word of the 24 x 24-bit product
    // (y3 2^10) m.  m = floor((2^22 - 1) / d) + 1 from a reciprocal estimate
    // (within one of the floor) corrected by its remainder.
    uint32_t f = (uint32_t)((float)0x3FFFFF * __builtin_amdgcn_rcpf((float)d));
    int r = 0x3FFFFF - (int)(f * d);
    if (r < 0) {
        f -= 1u;
        r += (int)d;
    }
    if (r >= (int)d) f += 1u;
    const uint32_t m = f + 1u;
    const uint32_t k = (4u + 3u * delta - 14u * mn) << 10;  // y3 2^10 = 14 v 2^10 + k (mod 2^32)
    // q lies in 1..10.  The selector order {1,7,6,5,4,3,2,0} of levels 0..7 is looked
    // up by v_perm_b32 from the bytes {0, 1, 7, 6, 5, 4, 3, 2} of the two table words
    // at index q; indices 8-10 give the sign byte of a positive byte, 0, level 7's
    // selector, and the result's other bytes index byte 0, 0.
    uint32_t lo = mx | (mn << 8), hi = 0;
#pragma unroll
    for (int i = 0; i < 16; i++) {
        const uint32_t q = mul_hi_u24(__umul24(v[i], 14u << 10) + k, m);
        const uint32_t sel = __builtin_amdgcn_perm(0x02030405u, 0x06070100u, q);
        const int b = 16 + 3 * i;  // the 3-bit selector of texel i at bit b of the block
        if (b < 32) lo |= sel << b;
        if (b > 29) hi |= b >= 32 ? sel << (b - 32) : sel >> (32 - b);
    }
    return make_uint2(lo, hi);
}

// The eight values of a BC4 block (quicktex/s3tc/bc4/BC4Block.cpp:44-64) as
// bytes, value s in byte s of (lo, hi): both rules computed, one kept, so lanes
// of either kind do not diverge.
__device__ __forceinline__ uint2 bc4_palette(uint32_t a0, uint32_t a1) {
    uint32_t six[4], eight[6];
#pragma unroll
    for (uint32_t s = 2; s < 6; s++) six[s - 2] = (a0 * (6 - s) + a1 * (s - 1)) / 5;
#pragma unroll
    for (uint32_t s = 2; s < 8; s++) eight[s - 2] = (a0 * (8 - s) + a1 * (s - 1)) / 7;
    const uint32_t base = a0 | (a1 << 8);
    return a0 <= a1 ? make_uint2(base | (six[0] << 16) | (six[1] << 24), six[2] | (six[3] << 8) | 0xFF000000u)
                    : make_uint2(base | (eight[0] << 16) | (eight[1] << 24),
                                 eight[2] | (eight[3] << 8) | (eight[4] << 16) | (eight[5] << 24));
}

// The texels' channel ch (byte sh / 8) set from a BC4 block (a0, a1, 48
// selector bits): each value picked out of the palette bytes by v_perm_b32 and
// placed at its byte in the same step (selector byte 0x0C gives zero).
__device__ __forceinline__ void bc4_decode_into(uint32_t px[16], uint32_t a0, uint32_t a1, uint64_t bits, uint32_t sh) {
    const uint2 pal = bc4_palette(a0, a1);
    const uint32_t keep = ~(0xFFu << sh), zero = 0x0C0C0C0Cu & keep;
#pragma unroll
    for (int i = 0; i < 16; i++) {
        const uint32_t sel = zero | (((uint32_t)(bits >> (3 * i)) & 7u) << sh);
        px[i] = (px[i] & keep) | __builtin_amdgcn_perm(pal.y, pal.x, sel);
    }
}

// ---------------------------------------------------------------------------
// One BC1 block (the colour half of BC3 as well, with its alpha block when bc3)
// per group of 16 lanes: gather the texels into the group's LDS words, encode,
// store.  Every lane of the group calls it with the same gid.
template <int INTERP>
__device__ __forceinline__ void encode_color_block(const uint8_t *__restrict__ rgba, int w, int h, int bw, int gid,
                                                   uint8_t *__restrict__ out, const DeviceTables *__restrict__ T,
                                                   bool bc3) {
    const int bx = gid % bw, by = gid / bw;
    const uint32_t px = load_texel(rgba, w, h, bx, by, glane());
    L().tex[threadIdx.x] = px;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const uint32_t rgb0 = px & 0x00FFFFFFu;
    const int pp = (int)gsum(dot4(rgb0, rgb0));
    const uint32_t ppa = gsum(dot4(px, px));
    const Ctx cx{T, px, pp, ppa};
    const uint2 colour = encode_bc1<INTERP>(cx);
    uint2 alpha = make_uint2(0, 0);
    if (bc3) {
        uint32_t a[16];
#pragma unroll
        for (int i = 0; i < 16; i++) a[i] = tex(i) >> 24;
        alpha = bc4_encode_values(a);
    }
    if (glane() == 0) {
        if (bc3) reinterpret_cast<uint4 *>(out)[gid] = make_uint4(alpha.x, alpha.y, colour.x, colour.y);
        else reinterpret_cast<uint2 *>(out)[gid] = colour;
    }
    // the group's next block overwrites these LDS words: keep the reads above them
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

template <int INTERP, bool BC3>
__global__ __launch_bounds__(256, 4) void k_bc1_encode(const uint8_t *__restrict__ rgba, int w, int h, int bw, int nblocks,
                                                       uint8_t *__restrict__ out, const DeviceTables *__restrict__ T,
                                                       BC1Settings s) {
    load_lds_tables<INTERP>(T, s);
    __syncthreads();
    // workgroups stride over chunks of 16 blocks; each group of 16 lanes owns one
    // block of the chunk and only touches its own 16 LDS texel words
    const int nchunks = (nblocks + 15) / 16;
    const int group = (int)(threadIdx.x / G);
    for (int chunk = (int)blockIdx.x; chunk < nchunks; chunk += (int)gridDim.x) {
        const int gid = chunk * 16 + group;
        if (gid >= nblocks) break;
        encode_color_block<INTERP>(rgba, w, h, bw, gid, out, T, BC3);
    }
}

// The 16 texels of block gid (index x + 4y) for one lane: four 16-byte row loads
// inside the texture, the wrap rule at its edges.
__device__ __forceinline__ void load_block16(const uint8_t *__restrict__ rgba, int w, int h, int bw, int gid,
                                             uint32_t (&px)[16]) {
    const int bx = gid % bw, by = gid / bw;
    if ((w & 3) == 0 && bx * 4 + 3 < w && by * 4 + 3 < h) {
#pragma unroll
        for (int y = 0; y < 4; y++) {
            const uint4 row = *reinterpret_cast<const uint4 *>(rgba + 4 * ((size_t)(by * 4 + y) * w + bx * 4));
            px[4 * y + 0] = row.x; px[4 * y + 1] = row.y; px[4 * y + 2] = row.z; px[4 * y + 3] = row.w;
        }
    } else {
#pragma unroll
        for (int t = 0; t < 16; t++) px[t] = load_texel(rgba, w, h, bx, by, t);
    }
}

// Stores block gid: the colour block, and for BC3 the alpha block of its texels
// before it.
__device__ __forceinline__ void store_lane_block(uint8_t *__restrict__ out, uint32_t gid, const uint32_t (&px)[16],
                                                 uint2 colour, bool bc3) {
    if (bc3) {
        uint32_t a[16];
#pragma unroll
        for (int i = 0; i < 16; i++) a[i] = px[i] >> 24;
        const uint2 alpha = bc4_encode_values(a);
        reinterpret_cast<uint4 *>(out)[gid] = make_uint4(alpha.x, alpha.y, colour.x, colour.y);
    } else {
        reinterpret_cast<uint2 *>(out)[gid] = colour;
    }
}

// One BC1 (or BC3) block per lane: encode and store.
template <int INTERP>
__device__ __forceinline__ void encode_lane_block(const uint8_t *__restrict__ rgba, int w, int h, int bw, int gid,
                                                  uint8_t *__restrict__ out, const DeviceTables *__restrict__ T,
                                                  const BC1Settings &s, bool bc3) {
    uint32_t px[16];
    load_block16(rgba, w, h, bw, gid, px);
    store_lane_block(out, (uint32_t)gid, px, lane::encode_bc1<INTERP, false>(px, s, T), bc3);
}

// BC1/BC3 encode, one block per lane (any settings; LIGHT: the settings of
// levels 0-12 as constants, lane::light_settings).
template <int INTERP, bool BC3, bool LIGHT>
__global__ __launch_bounds__(256, 4) void k_bc1_lane(const uint8_t *__restrict__ rgba, int w, int h, int bw, int nblocks,
                                                  uint8_t *__restrict__ out, const DeviceTables *__restrict__ T,
                                                  BC1Settings s_in) {
    const BC1Settings s = lane::fixed<LIGHT ? lane::kFixLight : lane::kFixNone>(s_in);
    lane::load_tables<INTERP, false>(T);
    __syncthreads();
    const int gid = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (gid < nblocks) encode_lane_block<INTERP>(rgba, w, h, bw, gid, out, T, s, BC3);
}

// queue[0] hands out units and queue[1] counts the waves that found it empty;
// the last of those waves returns both to zero, so the next launch of the same
// kernel on the same queue (stream-ordered after this one) starts from zero with
// no fill or memset in between.  Every wave calls this exactly once.
__device__ __forceinline__ void retire(uint32_t *queue, uint32_t waves) {
    if ((threadIdx.x & 63) == 0) {
        __threadfence();
        if (atomicAdd(&queue[1], 1u) == waves - 1) {
            atomicExch(&queue[0], 0u);
            atomicExch(&queue[1], 0u);
        }
    }
}

// The same encoder with the blocks dealt from queues, so that the searches run
// on full waves.  A solid block (one RGBA value: EncodeBlock writes it from the
// single-colour tables, BC1Encoder.cpp:262-268) costs a few dozen instructions,
// every other block tens of thousands; with one fixed block per lane, each solid
// block would leave its lane idle while the wave searches the others.  Here each
// lane holds a non-solid block before the wave starts a search: the lanes without
// one take the next block indices of the wave's current range (one atomic per
// wave and refill, indices ranked among the lanes that asked), store a solid block
// on the spot and ask again.
//
// The blocks are split into kLaneQueues contiguous ranges with one counter each
// (one address sustains only so many atomics per second); a wave starts on the
// range of its workgroup's XCD (workgroups go to the 8 XCDs round-robin) and moves
// on to the next range when its own is used up.  A wave claims exactly the blocks
// it places, so no blocks sit reserved by a wave when the ranges run dry: a search
// runs on fewer than 64 lanes only at the very end.  The grid is persistent (at
// most the resident workgroups).  The queue words (kLaneQueueWords, zero at
// launch) are left zero by the last wave to finish, so launches on one queue must
// be stream-ordered.
__device__ __forceinline__ void retire_lane_queue(uint32_t *queue, uint32_t waves) {
    if ((threadIdx.x & 63) == 0) {
        __threadfence();
        if (atomicAdd(&queue[kLaneRetireWord], 1u) == waves - 1) {
            for (int q = 0; q < kLaneQueues; q++) atomicExch(&queue[q * kLaneQueueStride], 0u);
            atomicExch(&queue[kLaneRetireWord], 0u);
        }
    }
}

// A parked search (kParkRecordBytes): block index, endpoints, error, selectors;
// next round, last improving round, mode | (forbidden move + 1) << 8 | improved << 16.
__device__ __forceinline__ void park_search(uint4 *__restrict__ park, uint32_t slot, uint32_t gid,
                                            const lane::SearchPark &pk) {
    park[2 * slot] = make_uint4(gid, pk.r.ep, pk.r.err, pk.r.sel);
    park[2 * slot + 1] = make_uint4(pk.st.next, pk.st.prev,
                                    (uint32_t)pk.r.mode | ((uint32_t)(pk.st.forbidden + 1) << 8) |
                                        ((uint32_t)pk.st.improved << 16),
                                    0u);
}

// F: the encoder build (lane::Fixed: any settings, those of levels 13-18 or those
// of levels 0-12 as constants).
template <int INTERP, bool BC3, int F>
__global__ __launch_bounds__(lane::kQueueThreads, 1) void k_bc1_lane_queue(const uint8_t *__restrict__ rgba, int w, int h, int bw,
                                                        uint32_t nblocks, uint8_t *__restrict__ out,
                                                        const DeviceTables *__restrict__ T, BC1Settings s_in,
                                                        uint32_t *__restrict__ queue, uint4 *__restrict__ park,
                                                        uint32_t park_stop) {
    constexpr uint32_t kNone = 0xFFFFFFFFu;
    const BC1Settings s = lane::fixed<F>(s_in);
    lane::load_tables<INTERP, true>(T);
    // the wave's current range and the number of ranges it found used up live in
    // LDS between refills, so that they hold no registers through the searches
    uint32_t *const ws = lane::LL<true>().wave_state[__builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6))];
    if ((threadIdx.x & 63) == 0) {
        ws[0] = blockIdx.x % kLaneQueues;
        ws[1] = 0;
    }
    __syncthreads();
    uint32_t px[16];
    bool open = true;  // wave-uniform: some range may still hold blocks
    for (;;) {
        uint32_t gid = kNone;
        if (open) {
            const uint32_t per = (nblocks + kLaneQueues - 1) / kLaneQueues;  // blocks per range
            uint32_t q = (uint32_t)__builtin_amdgcn_readfirstlane((int)ws[0]);
            uint32_t exhausted = (uint32_t)__builtin_amdgcn_readfirstlane((int)ws[1]);
            // refill until every lane holds a non-solid block or every range is used up
            for (;;) {
                const uint64_t need = __ballot(gid == kNone);
                if (need == 0) break;
                if (exhausted == kLaneQueues) {
                    open = false;
                    break;
                }
                const uint32_t want = (uint32_t)__popcll(need);
                uint32_t base = 0;
                if ((threadIdx.x & 63) == 0) base = atomicAdd(&queue[q * kLaneQueueStride], want);
                base = (uint32_t)__builtin_amdgcn_readfirstlane((int)base);
                const uint32_t begin = q * per, end = min(begin + per, nblocks);
                const uint32_t size = end > begin ? end - begin : 0u;
                const uint32_t got = base < size ? min(want, size - base) : 0u;
                if (gid == kNone) {
                    const uint32_t rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(need >> 32),
                                                                    __builtin_amdgcn_mbcnt_lo((uint32_t)need, 0u));
                    if (rank < got) {
                        const uint32_t idx = begin + base + rank;
                        load_block16(rgba, w, h, bw, (int)idx, px);
                        bool solid = true;
#pragma unroll
                        for (int i = 1; i < 16; i++) solid &= px[i] == px[0];
                        if (solid) store_lane_block(out, idx, px, lane::write_solid<true>(px[0]), BC3);
                        else gid = idx;
                    }
                }
                if (got < want) {  // this range is used up: continue on the next one
                    q = (q + 1) % kLaneQueues;
                    exhausted++;
                }
            }
            if ((threadIdx.x & 63) == 0) {
                ws[0] = q;
                ws[1] = exhausted;
            }
            // keep the state in LDS: no value of it is carried in registers past here
            asm volatile("" ::: "memory");
        }
        if (__ballot(gid != kNone) == 0) break;
        lane::SearchPark pk;
        pk.stop = park_stop;
        pk.parked = false;
        uint2 colour = make_uint2(0u, 0u);
        if (gid != kNone) colour = lane::encode_bc1<INTERP, true>(px, s, T, &pk);
        // every lane of the wave is active again here.  Searches with rounds left go
        // to the park buffer: one slot claim per wave; a block parks at most once,
        // so the slots never exceed nblocks.
        const uint64_t parked = __ballot(pk.parked);
        if (parked != 0) {
            uint32_t base = 0;
            if ((threadIdx.x & 63) == 0) base = atomicAdd(&queue[kParkCountWord], (uint32_t)__popcll(parked));
            base = (uint32_t)__builtin_amdgcn_readfirstlane((int)base);
            if (pk.parked) {
                const uint32_t rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(parked >> 32),
                                                                __builtin_amdgcn_mbcnt_lo((uint32_t)parked, 0u));
                park_search(park, base + rank, gid, pk);
            }
        }
        if (gid != kNone && !pk.parked) store_lane_block(out, gid, px, colour, BC3);
    }
    retire_lane_queue(queue, gridDim.x * (blockDim.x / 64));
}

// The drain gives each parked search a group of kDrainGroup lanes, which evaluate
// that many rounds of it at once (a speculative window), and deals the searches
// to the waves in equal contiguous shares (no atomics: one counter taking every
// claim would serialise them), at least kDrainPerGroup per group: enough waves to
// spread them over the CUs, few enough that a group whose search ends finds
// another to take.
constexpr int kDrainGroup = 8;
constexpr uint32_t kDrainGroups = 64 / kDrainGroup;  // groups per wave
constexpr uint32_t kDrainPerGroup = 2;
// A wave hands out new searches once this many of its groups are free.
constexpr uint32_t kDrainRefill = 2;

// The second half of the parked endpoint searches (k_bc1_lane_queue): each step
// runs one window of rounds of every busy group's search; finished searches get
// their selectors and are stored, and their groups take the next searches of the
// wave's share (in batches of kDrainRefill groups).  A wave stops when its share
// is used up and its groups are free: no wave waits for another.  The last wave
// to finish returns the park words to zero.
template <int INTERP, bool BC3>
__global__ __launch_bounds__(lane::kQueueThreads, 1) void k_bc1_search_drain(
    const uint8_t *__restrict__ rgba, int w, int h, int bw, uint8_t *__restrict__ out,
    const DeviceTables *__restrict__ T, BC1Settings s, const uint4 *__restrict__ park, uint32_t *__restrict__ queue) {
    lane::load_tables<INTERP, true>(T);
    __syncthreads();
    const uint32_t total = queue[kParkCountWord];  // final: counted by the previous launch
    const uint32_t wave = (uint32_t)__builtin_amdgcn_readfirstlane((int)(blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6)));
    const uint32_t waves = gridDim.x * (blockDim.x / 64);
    const uint32_t j = threadIdx.x % kDrainGroup;                 // lane in its group
    const uint32_t g = (threadIdx.x & 63u) / (uint32_t)kDrainGroup;  // group in the wave
    // this wave's share of the parked searches: [cursor, end)
    const uint32_t used = min(waves, (total + kDrainGroups * kDrainPerGroup - 1) / (kDrainGroups * kDrainPerGroup));
    const uint32_t share = used ? (total + used - 1) / used : 0u;
    uint32_t cursor = min(wave * share, total);
    const uint32_t end = min(cursor + share, total);
    if (cursor < end) {
        // the state of the group's search, the same in each of its lanes
        bool has = false, done = false, open = true;
        uint32_t gid = 0, px[16];
        lane::Blk bk{0u, 0u};
        Res r{0u, 0u, 0u, 0};
        lane::SearchState st{0u, 0u, -1, false};
        for (;;) {
            const uint32_t busy = (uint32_t)__popcll(__ballot(has && !done)) / (uint32_t)kDrainGroup;
            if (kDrainGroups - busy >= kDrainRefill || busy == 0) {
                if (done) {  // finish: selectors, then the block (BC3: its alpha from the source)
                    lane::search_finish<INTERP, true>(px, bk, s.error_mode, r, st);
                    const uint2 colour = write_block(r);
                    if (j == 0) {
                        if (BC3) load_block16(rgba, w, h, bw, (int)gid, px);
                        store_lane_block(out, gid, px, colour, BC3);
                    }
                    has = false;
                    done = false;
                }
                if (open) {
                    const uint64_t want = __ballot(!has && j == 0);  // the free groups' first lanes
                    const uint32_t nwant = (uint32_t)__popcll(want);
                    const uint32_t base = cursor;
                    const uint32_t got = min(nwant, end - base);
                    cursor += got;
                    if (cursor == end) open = false;
                    if (!has) {
                        const uint32_t rank = (uint32_t)__popcll(want & ((1ull << (g * kDrainGroup)) - 1ull));
                        if (rank < got) {
                            const uint4 a = park[2 * (base + rank)], b = park[2 * (base + rank) + 1];
                            gid = a.x;
                            r = Res{a.y, a.w, a.z, (int)(b.z & 0xFFu)};
                            st = lane::SearchState{b.x, b.y, (int)((b.z >> 8) & 0xFFu) - 1, ((b.z >> 16) & 1u) != 0};
                            load_block16(rgba, w, h, bw, (int)gid, px);
                            bk.pp = 0u;
#pragma unroll
                            for (int i = 0; i < 16; i++) {
                                px[i] &= 0x00FFFFFFu;
                                bk.pp += dot4(px[i], px[i]);
                            }
                            has = true;
                        }
                    }
                }
                if (__ballot(has) == 0) break;
            }
            if (has && !done)
                done = lane::search_window<INTERP, true, kDrainGroup>(px, bk, s.search_rounds, s.error_mode, r, st);
        }
    }
    if ((threadIdx.x & 63) == 0) {
        __threadfence();
        if (atomicAdd(&queue[kParkRetireWord], 1u) == gridDim.x * (blockDim.x / 64) - 1) {
            atomicExch(&queue[kParkCountWord], 0u);
            atomicExch(&queue[kParkRetireWord], 0u);
        }
    }
}

// The 16 texels of block gid (edge blocks wrap like the reference's texel reads).
__device__ __forceinline__ void load_block_texels(const uint8_t *__restrict__ rgba, int w, int h, int bw, int gid,
                                                  uint32_t px[16]) {
    const int bx = gid % bw, by = gid / bw;
    if ((w & 3) == 0 && bx * 4 + 3 < w && by * 4 + 3 < h) {
        // every texel is read once: non-temporal loads keep it out of the caches
        typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
#pragma unroll
        for (int y = 0; y < 4; y++) {
            const u32x4 row = __builtin_nontemporal_load(
                reinterpret_cast<const u32x4 *>(rgba + 4 * ((size_t)(by * 4 + y) * w + bx * 4)));
            px[4 * y + 0] = row.x; px[4 * y + 1] = row.y; px[4 * y + 2] = row.z; px[4 * y + 3] = row.w;
        }
    } else {
#pragma unroll
        for (int t = 0; t < 16; t++) px[t] = load_texel(rgba, w, h, bx, by, t);
    }
}

// BC4 (one channel) or BC5 (two channels) of block gid: one lane per block.
template <bool BC5>
__device__ __forceinline__ void encode_bc45_block(const uint8_t *__restrict__ rgba, int w, int h, int bw, int gid, int ch0,
                                                  int ch1, uint8_t *__restrict__ out) {
    uint32_t px[16];
    load_block_texels(rgba, w, h, bw, gid, px);
    uint32_t v[16];
#pragma unroll
    for (int i = 0; i < 16; i++) v[i] = chn(px[i], ch0);
    const uint2 b0 = bc4_encode_values(v);
    if (BC5) {
#pragma unroll
        for (int i = 0; i < 16; i++) v[i] = chn(px[i], ch1);
        const uint2 b1 = bc4_encode_values(v);
        reinterpret_cast<uint4 *>(out)[gid] = make_uint4(b0.x, b0.y, b1.x, b1.y);
    } else {
        reinterpret_cast<uint2 *>(out)[gid] = b0;
    }
}

template <bool BC5>
__global__ __launch_bounds__(256) void k_bc4_encode(const uint8_t *__restrict__ rgba, int w, int h, int bw, int nblocks,
                                                    int ch0, int ch1, uint8_t *__restrict__ out) {
    const int gid = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (gid < nblocks) encode_bc45_block<BC5>(rgba, w, h, bw, gid, ch0, ch1, out);
}

// ---------------------------------------------------------------------------
// Decoders: one lane per block.
enum DecKind { DEC_BC1 = 1, DEC_BC3 = 3, DEC_BC4 = 4, DEC_BC5 = 5 };

// true if texel t of block b is the final writer of its pixel along an axis of
// n pixels and nb blocks.
__device__ __forceinline__ bool axis_winner(int b, int t, int n, int nb) {
    const int X = (b * 4 + t) % n;
    const int last = nb - 1;
#pragma unroll
    for (int tt = 3; tt >= 0; tt--) {
        if ((last * 4 + tt) % n == X) return b == last && t == tt;
    }
    return b == (X >> 2) && t == (X & 3);
}

// A 16-byte row of decoded texels, stored non-temporally: the decoders write 8x
// the bytes they read and nothing reads the texels back, so the stores stream
// past the caches (measured 2.1x faster on a 4096x4096 decode than plain stores).
__device__ __forceinline__ void store_row_nt(uint8_t *row, uint32_t a, uint32_t b, uint32_t c, uint32_t d) {
    typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
    const u32x4 v = {a, b, c, d};
    __builtin_nontemporal_store(v, reinterpret_cast<u32x4 *>(row));
}

// Decode of block gid of a bw x bh block texture into the w x h RGBA8 texture.
template <int INTERP, int KIND>
__device__ __forceinline__ void decode_block(const uint8_t *__restrict__ blocks, int w, int h, int bw, int bh, int p0,
                                             int p1, uint8_t *__restrict__ rgba, int gid) {
    const int bx = gid % bw, by = gid / bw;
    uint32_t px[16];
#pragma unroll
    for (int i = 0; i < 16; i++) px[i] = 0xFF000000u;  // Color() = (0, 0, 0, 255)
    if (KIND == DEC_BC1 || KIND == DEC_BC3) {
        uint2 cb;
        uint2 ab = make_uint2(0, 0);
        if (KIND == DEC_BC3) {
            const uint4 q = reinterpret_cast<const uint4 *>(blocks)[gid];
            ab = make_uint2(q.x, q.y);
            cb = make_uint2(q.z, q.w);
        } else {
            cb = reinterpret_cast<const uint2 *>(blocks)[gid];
        }
        // BC1Decoder::DecodeBlock (BC1Decoder.cpp:31-52) via Interpolate565BC1 (Interpolator.cpp:52-55)
        const uint32_t c0 = cb.x & 0xFFFFu, c1 = cb.x >> 16;
        const bool three = (KIND == DEC_BC1) && (c1 >= c0);
        const uint32_t lo = rgb(c0 >> 11, (c0 >> 5) & 63u, c0 & 31u);
        const uint32_t hi = rgb(c1 >> 11, (c1 >> 5) & 63u, c1 & 31u);
        uint32_t pal[4];
        palette<INTERP>(lo, hi, three, pal);
        // back to BC1 selector order {c0, c1, c2, c3}; alpha 255 except 3-colour black
        // (3-colour black keeps alpha 0 only when alpha is written; BC3 takes alpha
        // from its alpha block below)
        const bool write_alpha = (KIND == DEC_BC1) && p0 != 0;
        const uint32_t bc0 = pal[0] | 0xFF000000u;
        const uint32_t bc1 = (three ? pal[2] : pal[3]) | 0xFF000000u;
        const uint32_t bc2 = pal[1] | 0xFF000000u;
        const uint32_t bc3 = three ? (write_alpha ? 0u : 0xFF000000u) : (pal[2] | 0xFF000000u);
        if (KIND == DEC_BC1) {
            // byte planes: R holds the red of colours 0-3 in bytes 0-3, and so on; a
            // texel is then two v_perm_b32 (red and green, blue and alpha) of its selector
            const uint32_t lo01 = __builtin_amdgcn_perm(bc1, bc0, 0x05010400u);
            const uint32_t hi01 = __builtin_amdgcn_perm(bc1, bc0, 0x07030602u);
            const uint32_t lo23 = __builtin_amdgcn_perm(bc3, bc2, 0x05010400u);
            const uint32_t hi23 = __builtin_amdgcn_perm(bc3, bc2, 0x07030602u);
            const uint32_t R = __builtin_amdgcn_perm(lo23, lo01, 0x05040100u);
            const uint32_t G = __builtin_amdgcn_perm(lo23, lo01, 0x07060302u);
            const uint32_t B = __builtin_amdgcn_perm(hi23, hi01, 0x05040100u);
            const uint32_t A = __builtin_amdgcn_perm(hi23, hi01, 0x07060302u);
#pragma unroll
            for (int i = 0; i < 16; i++) {
                const uint32_t s = (cb.y >> (2 * i)) & 3u;
                const uint32_t sel = s * 0x101u + 0x0C0C0400u;  // bytes: s, s + 4, zero, zero
                px[i] = __builtin_amdgcn_perm(G, R, sel) | __builtin_amdgcn_perm(A, B, (sel << 16) | 0x0C0Cu);
            }
        } else {
#pragma unroll
            for (int i = 0; i < 16; i++) {
                const uint32_t s = (cb.y >> (2 * i)) & 3u;
                px[i] = s == 0 ? bc0 : (s == 1 ? bc1 : (s == 2 ? bc2 : bc3));
            }
        }
        if (KIND == DEC_BC3)
            bc4_decode_into(px, ab.x & 0xFFu, (ab.x >> 8) & 0xFFu, (uint64_t)(ab.x >> 16) | ((uint64_t)ab.y << 16), 24u);
    } else {
        const int nch = KIND == DEC_BC5 ? 2 : 1;
#pragma unroll
        for (int k = 0; k < nch; k++) {
            const uint2 q = reinterpret_cast<const uint2 *>(blocks)[gid * nch + k];
            const int ch = k == 0 ? p0 : p1;
            bc4_decode_into(px, q.x & 0xFFu, (q.x >> 8) & 0xFFu, (uint64_t)(q.x >> 16) | ((uint64_t)q.y << 16),
                            8u * (uint32_t)ch);
        }
    }
    // scatter
    const int wrapx = bw * 4 - w, wrapy = bh * 4 - h;
    const bool fast_x = wrapx == 0 || (bx != 0 && bx != bw - 1);
    const bool fast_y = wrapy == 0 || (by != 0 && by != bh - 1);
    if (fast_x && fast_y) {
#pragma unroll
        for (int y = 0; y < 4; y++) {
            uint8_t *row = rgba + 4 * ((size_t)(by * 4 + y) * w + bx * 4);
            if ((w & 3) == 0) {
                store_row_nt(row, px[4 * y], px[4 * y + 1], px[4 * y + 2], px[4 * y + 3]);
            } else {
#pragma unroll
                for (int x = 0; x < 4; x++) reinterpret_cast<uint32_t *>(row)[x] = px[4 * y + x];
            }
        }
    } else {
        bool wx[4], wy[4];
#pragma unroll
        for (int t = 0; t < 4; t++) {
            wx[t] = axis_winner(bx, t, w, bw);
            wy[t] = axis_winner(by, t, h, bh);
        }
#pragma unroll
        for (int y = 0; y < 4; y++)
#pragma unroll
            for (int x = 0; x < 4; x++) {
                if (!(wx[x] && wy[y])) continue;
                const int X = (bx * 4 + x) % w, Y = (by * 4 + y) % h;
                *reinterpret_cast<uint32_t *>(rgba + 4 * ((size_t)Y * w + X)) = px[4 * y + x];
            }
    }
}

template <int INTERP, int KIND>
__global__ __launch_bounds__(256) void k_decode(const uint8_t *__restrict__ blocks, int w, int h, int bw, int bh,
                                                int p0, int p1, uint8_t *__restrict__ rgba) {
    const int gid = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (gid < bw * bh) decode_block<INTERP, KIND>(blocks, w, h, bw, bh, p0, p1, rgba, gid);
}

// ---------------------------------------------------------------------------
// Batch kernels.  A batch is a list of jobs (textures); each kernel of a batch
// owns a job array sorted by descending cost and a queue of work units numbered
// job by job.

// Job of unit u: the last job whose unit_begin <= u, searched upward from a
// job at or before it (units are taken in increasing order, so a wave's next
// unit is in its current job or a later one).
__device__ __forceinline__ int find_job(const BatchJob *__restrict__ jobs, int from, int njobs, uint32_t u) {
    if (from + 1 >= njobs || jobs[from + 1].unit_begin > u) return from;
    int lo = from + 1, hi = njobs - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (jobs[mid].unit_begin <= u) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// Next unit of the launch's queue for the calling wave (one atomic per wave).
__device__ __forceinline__ uint32_t take_unit(uint32_t *queue) {
    uint32_t u = 0;
    if ((threadIdx.x & 63) == 0) u = atomicAdd(&queue[0], 1u);
    return (uint32_t)__builtin_amdgcn_readfirstlane((int)u);
}


// Persistent BC1/BC3 encoder over the jobs of one interpolator.  Each wave takes
// a unit (unit_blocks consecutive blocks of one job), loads the job's settings
// into its LDS slot when the job changes, and encodes the unit four blocks at a
// time (one per group).
template <int INTERP>
__global__ __launch_bounds__(256, 4) void k_color_batch(const BatchJob *__restrict__ jobs, int njobs, uint32_t nunits,
                                                        uint32_t *__restrict__ queue,
                                                        const DeviceTables *__restrict__ T) {
    load_lds_tables<INTERP>(T, jobs[0].s);
    __syncthreads();
    const int lane = (int)(threadIdx.x & 63);
    const int slot = wave_slot();
    const uint32_t sub = (threadIdx.x & 63) >> 4;  // group within the wave
    int cur = 0;
    for (;;) {
        const uint32_t u = take_unit(queue);
        if (u >= nunits) break;
        const int j = find_job(jobs, cur, njobs, u);
        if (j != cur) {
            // all reads of the slot by this wave were issued before this point, and
            // one wave's LDS operations complete in order
            if (lane < (int)(sizeof(BC1Settings) / 4))
                reinterpret_cast<uint32_t *>(&L().s[slot])[lane] = reinterpret_cast<const uint32_t *>(&jobs[j].s)[lane];
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            cur = j;
        }
        const BatchJob &J = jobs[j];
        const uint32_t b0 = (u - J.unit_begin) * J.unit_blocks;
        const uint32_t b1 = min(b0 + J.unit_blocks, J.nblocks);
        const bool bc3 = J.kind == BJ_ENC_BC3;
        for (uint32_t b = b0; b < b1; b += 4) {
            const uint32_t gid = b + sub;
            if (gid < b1) encode_color_block<INTERP>(J.src, J.w, J.h, J.bw, (int)gid, J.dst, T, bc3);
        }
    }
    retire(queue, gridDim.x * kWaves);
}

// Job of the workgroup's unit: binary search on unit_begin.
__device__ __forceinline__ const BatchJob &unit_job(const BatchJob *__restrict__ jobs, int njobs, uint32_t u) {
    int lo = 0, hi = njobs - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (jobs[mid].unit_begin <= u) lo = mid;
        else hi = mid - 1;
    }
    return jobs[lo];
}

// BC1/BC3 jobs without endpoint search of one interpolator: one block per lane,
// units of 256 blocks; workgroups stride over the units (one unit each when the
// grid has a workgroup per unit).
template <int INTERP>
__global__ __launch_bounds__(256, 4) void k_lane_batch(const BatchJob *__restrict__ jobs, int njobs, uint32_t nunits,
                                                    const DeviceTables *__restrict__ T) {
    lane::load_tables<INTERP, false>(T);
    __syncthreads();
    for (uint32_t u = blockIdx.x; u < nunits; u += gridDim.x) {
        const BatchJob &J = unit_job(jobs, njobs, u);
        const uint32_t gid = (u - J.unit_begin) * 256u + threadIdx.x;
        if (gid < J.nblocks)
            encode_lane_block<INTERP>(J.src, J.w, J.h, J.bw, (int)gid, J.dst, T, J.s, J.kind == BJ_ENC_BC3);
    }
}

// BC4/BC5 encode and every decoder of a batch: one lane per block, one unit of
// 256 blocks per workgroup.
template <int KIND>
__device__ __forceinline__ void decode_interp(const BatchJob &J, int gid) {
    switch (J.interp) {
        case QTX_INTERP_IDEAL_ROUND:
            decode_block<QTX_INTERP_IDEAL_ROUND, KIND>(J.src, J.w, J.h, J.bw, J.bh, J.p0, J.p1, J.dst, gid);
            break;
        case QTX_INTERP_NVIDIA:
            decode_block<QTX_INTERP_NVIDIA, KIND>(J.src, J.w, J.h, J.bw, J.bh, J.p0, J.p1, J.dst, gid);
            break;
        case QTX_INTERP_AMD: decode_block<QTX_INTERP_AMD, KIND>(J.src, J.w, J.h, J.bw, J.bh, J.p0, J.p1, J.dst, gid); break;
        default: decode_block<QTX_INTERP_IDEAL, KIND>(J.src, J.w, J.h, J.bw, J.bh, J.p0, J.p1, J.dst, gid); break;
    }
}

// Workgroups stride over the units, so a launch may use fewer workgroups than
// units (a decode that writes host memory is bound by PCIe, not by the CUs).
__global__ __launch_bounds__(256) void k_light_batch(const BatchJob *__restrict__ jobs, int njobs, uint32_t nunits) {
    for (uint32_t u = blockIdx.x; u < nunits; u += gridDim.x) {
        const BatchJob &J = unit_job(jobs, njobs, u);
        const uint32_t gid = (u - J.unit_begin) * 256u + threadIdx.x;
        if (gid >= J.nblocks) continue;
        switch (J.kind) {
            case BJ_ENC_BC4: encode_bc45_block<false>(J.src, J.w, J.h, J.bw, (int)gid, J.p0, J.p1, J.dst); break;
            case BJ_ENC_BC5: encode_bc45_block<true>(J.src, J.w, J.h, J.bw, (int)gid, J.p0, J.p1, J.dst); break;
            case BJ_DEC_BC1: decode_interp<DEC_BC1>(J, (int)gid); break;
            case BJ_DEC_BC3: decode_interp<DEC_BC3>(J, (int)gid); break;
            case BJ_DEC_BC4: decode_block<0, DEC_BC4>(J.src, J.w, J.h, J.bw, J.bh, J.p0, J.p1, J.dst, (int)gid); break;
            case BJ_DEC_BC5: decode_block<0, DEC_BC5>(J.src, J.w, J.h, J.bw, J.bh, J.p0, J.p1, J.dst, (int)gid); break;
            default: break;
        }
    }
}

// Writes zero to n words (a kernel, not a memset node, so that a caller's graph
// capture of the launch replays it faithfully).
__global__ __launch_bounds__(256) void k_zero_words(uint32_t *__restrict__ words, uint32_t n) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) words[i] = 0u;
}

}  // namespace dev

// ---------------------------------------------------------------------------
// Launchers
namespace {
constexpr int kThreads = 256;
constexpr size_t kBc1Lds = sizeof(dev::Lds);
// the lane kernels' LDS is static (dev::lane::LaneLdsPal / LaneLdsPlain): no dynamic bytes
constexpr size_t kLaneLds = 0, kLaneQueueLds = 0, kDrainLds = 0;
// persistent workgroups for the BC1/BC3 encoder: 256 CUs x 4 resident workgroups
constexpr int kBc1MaxGrid = 256 * 4;
inline unsigned grid_for(long long items, int per_block) { return (unsigned)((items + per_block - 1) / per_block); }
}  // namespace

// A fit of the oracle's op accounting on the block mixture (levels 0-19): about 2k
// ops of metrics, endpoints and selectors, 500 per cluster-fit ordering and pass,
// and 780 per endpoint-search round, of which about 37 run before the search stops
// for lack of improvement.
double bc1_block_ops(const BC1Settings &s) {
    const double o4 = s.exhaustive ? kOrders4 : s.orderings4;
    const double o3 = s.exhaustive ? kOrders3 : s.orderings3;
    const bool three = (s.color_mode & QTX_COLOR_THREE) != 0;
    double ops = 2000.0 + s.cf_passes * (o4 * 500.0 + (three ? o3 * 500.0 + 1000.0 : 0.0));
    ops += 780.0 * (s.search_rounds < 37u ? (double)s.search_rounds : 37.0);
    return ops;
}

// Crossovers measured on MI355X (ThreeColorBlack, block mixture, Ideal): the static
// lane kernel (levels 0-9) wins from 32k blocks; the queued kernel, whose waves run
// one whole block per lane (a fixed cost of one block's latency), from about
// 65k + ops/2 blocks: 81k at level 10, 93k at level 13, 125k at level 16, 150k at
// level 18.  Level 19 (exhaustive cluster fit, about 1 M ops per block): the queued
// kernel takes 3.7 ms at any size up to 131k blocks (one block per lane) against
// 0.23-3.0 ms for the 16-lane encoder, 4.8 against 5.9 ms at 262k blocks and 12.0
// against 20.5 ms at 1 M: from kLaneFullBlocks (the formula gives 565k).
bool use_lane_encoder(const BC1Settings &s, uint64_t blocks, int path) {
    const int p = path != QTX_BC1_AUTO ? path : bc1_path_env();
    if (p == QTX_BC1_GROUP) return false;
    if (p == QTX_BC1_LANE) return true;
    if (blocks >= kLaneFullBlocks) return true;
    const double ops = bc1_block_ops(s);
    if (ops < kLaneCheapOps) return blocks >= kLaneMinBlocks;
    return (double)blocks >= kLaneQueueBaseBlocks + 0.5 * ops;
}

// The batch's lane kernel (k_lane_batch) deals no blocks and parks no searches:
// it wins from 32k blocks at settings up to 120 k ops/block, from 262k above.
bool use_lane_batch(const BC1Settings &s, uint64_t blocks, int path) {
    const int p = path != QTX_BC1_AUTO ? path : bc1_path_env();
    if (p == QTX_BC1_GROUP) return false;
    if (p == QTX_BC1_LANE) return dev::lane::lane_settings(s.exhaustive);
    if (!dev::lane::lane_settings(s.exhaustive)) return false;
    if (blocks >= kLaneFullBlocks) return true;
    return blocks >= kLaneMinBlocks && bc1_block_ops(s) <= kLaneBatchMaxOps;
}

bool wants_lane_queue(const BC1Settings &s, int w, int h, int path) {
    const uint64_t nblocks = (uint64_t)((w + 3) / 4) * (uint64_t)((h + 3) / 4);
    return use_lane_encoder(s, nblocks, path) && lane_queue_env() && bc1_block_ops(s) >= kLaneQueueMinOps;
}

size_t bc1_park_bytes(const BC1Settings &s, int w, int h, int path) {
    const uint32_t stop = park_stop_env();
    if (!wants_lane_queue(s, w, h, path) || stop == 0 || s.search_rounds <= stop) return 0;
    return kParkRecordBytes * (size_t)((w + 3) / 4) * (size_t)((h + 3) / 4);
}

hipError_t launch_bc1_encode(const uint8_t *d_rgba, int w, int h, uint8_t *d_out, const DeviceTables *d_tables,
                             const BC1Settings &s, int interp, bool bc3, int path, uint32_t *d_queue, void *d_park,
                             int cu_count, hipStream_t stream) {
    const int bw = (w + 3) / 4, bh = (h + 3) / 4;
    const int nblocks = bw * bh;
    if (use_lane_encoder(s, (uint64_t)nblocks, path)) {
        const dim3 lblock(kThreads);
        if (d_queue && wants_lane_queue(s, w, h, path)) {
            // persistent: one workgroup of 12 waves per CU, fewer for small launches
            const dim3 qblock(dev::lane::kQueueThreads);
            const dim3 qgrid((unsigned)std::min<long long>(grid_for(nblocks, dev::lane::kQueueThreads),
                                                           std::max(cu_count, 1)));
            // parking (and the drain launch after the queued kernel) only with a park buffer
            const bool drain = d_park && bc1_park_bytes(s, w, h, path) > 0;
            const uint32_t stop = drain ? park_stop_env() : 0xFFFFFFFFu;
            uint4 *const park = static_cast<uint4 *>(d_park);
            const dim3 dgrid((unsigned)std::max(cu_count, 1));
            const int fix = dev::lane::fixed_build(s);
#define QTX_QUEUE_LAUNCH(I, B, H)                                                                                   \
    hipLaunchKernelGGL((dev::k_bc1_lane_queue<I, B, H>), qgrid, qblock, kLaneQueueLds, stream, d_rgba, w, h, bw,    \
                       (uint32_t)nblocks, d_out, d_tables, s, d_queue, park, stop)
#define QTX_LANE_CASE(I)                                                                                            \
    case I:                                                                                                         \
        if (bc3) {                                                                                                  \
            if (fix == dev::lane::kFixHQ) QTX_QUEUE_LAUNCH(I, true, dev::lane::kFixHQ);                             \
            else if (fix == dev::lane::kFixLight) QTX_QUEUE_LAUNCH(I, true, dev::lane::kFixLight);                  \
            else QTX_QUEUE_LAUNCH(I, true, dev::lane::kFixNone);                                                    \
            if (drain)                                                                                              \
                hipLaunchKernelGGL((dev::k_bc1_search_drain<I, true>), dgrid, qblock, kDrainLds, stream, d_rgba, w, \
                                   h, bw, d_out, d_tables, s, park, d_queue);                                       \
        } else {                                                                                                    \
            if (fix == dev::lane::kFixHQ) QTX_QUEUE_LAUNCH(I, false, dev::lane::kFixHQ);                            \
            else if (fix == dev::lane::kFixLight) QTX_QUEUE_LAUNCH(I, false, dev::lane::kFixLight);                 \
            else QTX_QUEUE_LAUNCH(I, false, dev::lane::kFixNone);                                                   \
            if (drain)                                                                                              \
                hipLaunchKernelGGL((dev::k_bc1_search_drain<I, false>), dgrid, qblock, kDrainLds, stream, d_rgba, w,\
                                   h, bw, d_out, d_tables, s, park, d_queue);                                       \
        }                                                                                                           \
        break;
            switch (interp) {
                QTX_LANE_CASE(QTX_INTERP_IDEAL)
                QTX_LANE_CASE(QTX_INTERP_IDEAL_ROUND)
                QTX_LANE_CASE(QTX_INTERP_NVIDIA)
                QTX_LANE_CASE(QTX_INTERP_AMD)
                default: return hipErrorInvalidValue;
            }
#undef QTX_LANE_CASE
#undef QTX_QUEUE_LAUNCH
            return hipGetLastError();
        }
        const dim3 lgrid(grid_for(nblocks, kThreads));
        const bool light = dev::lane::light_settings(s);
#define QTX_LANE_LAUNCH(I, B, L)                                                                                    \
    hipLaunchKernelGGL((dev::k_bc1_lane<I, B, L>), lgrid, lblock, kLaneLds, stream, d_rgba, w, h, bw, nblocks, d_out, \
                       d_tables, s)
#define QTX_LANE_CASE(I)                                                                                            \
    case I:                                                                                                         \
        if (bc3) {                                                                                                  \
            if (light) QTX_LANE_LAUNCH(I, true, true);                                                              \
            else QTX_LANE_LAUNCH(I, true, false);                                                                   \
        } else {                                                                                                    \
            if (light) QTX_LANE_LAUNCH(I, false, true);                                                             \
            else QTX_LANE_LAUNCH(I, false, false);                                                                  \
        }                                                                                                           \
        break;
        switch (interp) {
            QTX_LANE_CASE(QTX_INTERP_IDEAL)
            QTX_LANE_CASE(QTX_INTERP_IDEAL_ROUND)
            QTX_LANE_CASE(QTX_INTERP_NVIDIA)
            QTX_LANE_CASE(QTX_INTERP_AMD)
            default: return hipErrorInvalidValue;
        }
#undef QTX_LANE_CASE
#undef QTX_LANE_LAUNCH
        return hipGetLastError();
    }
    const dim3 grid((unsigned)std::min<long long>(((long long)nblocks + 15) / 16, kBc1MaxGrid)), block(kThreads);
#define QTX_BC1_CASE(I)                                                                                             \
    case I:                                                                                                         \
        if (bc3)                                                                                                    \
            hipLaunchKernelGGL((dev::k_bc1_encode<I, true>), grid, block, kBc1Lds, stream, d_rgba, w, h, bw, nblocks, d_out, \
                               d_tables, s);                                                                        \
        else                                                                                                        \
            hipLaunchKernelGGL((dev::k_bc1_encode<I, false>), grid, block, kBc1Lds, stream, d_rgba, w, h, bw, nblocks,      \
                               d_out, d_tables, s);                                                                 \
        break;
    switch (interp) {
        QTX_BC1_CASE(QTX_INTERP_IDEAL)
        QTX_BC1_CASE(QTX_INTERP_IDEAL_ROUND)
        QTX_BC1_CASE(QTX_INTERP_NVIDIA)
        QTX_BC1_CASE(QTX_INTERP_AMD)
        default: return hipErrorInvalidValue;
    }
#undef QTX_BC1_CASE
    return hipGetLastError();
}

hipError_t launch_zero_words(uint32_t *d_words, uint32_t n, hipStream_t stream) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(dev::k_zero_words, dim3(std::min<unsigned>(grid_for(n, kThreads), 64u)), dim3(kThreads), 0,
                       stream, d_words, n);
    return hipGetLastError();
}

hipError_t launch_bc4_encode(const uint8_t *d_rgba, int w, int h, uint8_t *d_out, int ch0, int ch1, bool bc5,
                             hipStream_t stream) {
    const int bw = (w + 3) / 4, bh = (h + 3) / 4;
    const int nblocks = bw * bh;
    const dim3 grid(grid_for(nblocks, kThreads)), block(kThreads);
    if (bc5)
        hipLaunchKernelGGL((dev::k_bc4_encode<true>), grid, block, 0, stream, d_rgba, w, h, bw, nblocks, ch0, ch1, d_out);
    else
        hipLaunchKernelGGL((dev::k_bc4_encode<false>), grid, block, 0, stream, d_rgba, w, h, bw, nblocks, ch0, ch1, d_out);
    return hipGetLastError();
}

hipError_t launch_decode(int format, const uint8_t *d_blocks, int w, int h, uint8_t *d_rgba, int interp, int p0, int p1,
                         hipStream_t stream) {
    const int bw = (w + 3) / 4, bh = (h + 3) / 4;
    const dim3 grid(grid_for((long long)bw * bh, kThreads)), block(kThreads);
#define QTX_DEC(I, K) \
    hipLaunchKernelGGL((dev::k_decode<I, K>), grid, block, 0, stream, d_blocks, w, h, bw, bh, p0, p1, d_rgba)
    switch (format) {
        case QTX_FMT_BC4: QTX_DEC(0, dev::DEC_BC4); break;
        case QTX_FMT_BC5: QTX_DEC(0, dev::DEC_BC5); break;
        case QTX_FMT_BC1:
        case QTX_FMT_BC3: {
            const bool b3 = format == QTX_FMT_BC3;
            switch (interp) {
                case QTX_INTERP_IDEAL: if (b3) QTX_DEC(QTX_INTERP_IDEAL, dev::DEC_BC3); else QTX_DEC(QTX_INTERP_IDEAL, dev::DEC_BC1); break;
                case QTX_INTERP_IDEAL_ROUND: if (b3) QTX_DEC(QTX_INTERP_IDEAL_ROUND, dev::DEC_BC3); else QTX_DEC(QTX_INTERP_IDEAL_ROUND, dev::DEC_BC1); break;
                case QTX_INTERP_NVIDIA: if (b3) QTX_DEC(QTX_INTERP_NVIDIA, dev::DEC_BC3); else QTX_DEC(QTX_INTERP_NVIDIA, dev::DEC_BC1); break;
                case QTX_INTERP_AMD: if (b3) QTX_DEC(QTX_INTERP_AMD, dev::DEC_BC3); else QTX_DEC(QTX_INTERP_AMD, dev::DEC_BC1); break;
                default: return hipErrorInvalidValue;
            }
            break;
        }
        default: return hipErrorInvalidValue;
    }
#undef QTX_DEC
    return hipGetLastError();
}

hipError_t launch_color_batch(int interp, const BatchJob *d_jobs, int njobs, uint32_t nunits, uint32_t *d_queue,
                              unsigned grid, const DeviceTables *d_tables, hipStream_t stream) {
    const dim3 block(kThreads);
#define QTX_CB(I)                                                                                                   \
    case I:                                                                                                         \
        hipLaunchKernelGGL((dev::k_color_batch<I>), dim3(grid), block, kBc1Lds, stream, d_jobs, njobs, nunits, d_queue, \
                           d_tables);                                                                               \
        break;
    switch (interp) {
        QTX_CB(QTX_INTERP_IDEAL)
        QTX_CB(QTX_INTERP_IDEAL_ROUND)
        QTX_CB(QTX_INTERP_NVIDIA)
        QTX_CB(QTX_INTERP_AMD)
        default: return hipErrorInvalidValue;
    }
#undef QTX_CB
    return hipGetLastError();
}

hipError_t launch_lane_batch(int interp, const BatchJob *d_jobs, int njobs, uint32_t nunits, unsigned grid,
                             const DeviceTables *d_tables, hipStream_t stream) {
    const dim3 g(std::max(1u, std::min(grid, nunits))), block(kThreads);
#define QTX_LB(I)                                                                                                   \
    case I:                                                                                                         \
        hipLaunchKernelGGL((dev::k_lane_batch<I>), g, block, kLaneLds, stream, d_jobs, njobs, nunits, d_tables);    \
        break;
    switch (interp) {
        QTX_LB(QTX_INTERP_IDEAL)
        QTX_LB(QTX_INTERP_IDEAL_ROUND)
        QTX_LB(QTX_INTERP_NVIDIA)
        QTX_LB(QTX_INTERP_AMD)
        default: return hipErrorInvalidValue;
    }
#undef QTX_LB
    return hipGetLastError();
}

hipError_t launch_light_batch(const BatchJob *d_jobs, int njobs, uint32_t nunits, unsigned grid, hipStream_t stream) {
    hipLaunchKernelGGL(dev::k_light_batch, dim3(std::max(1u, std::min(grid, nunits))), dim3(kThreads), 0, stream, d_jobs,
                       njobs, nunits);
    return hipGetLastError();
}

}  // namespace qtx
