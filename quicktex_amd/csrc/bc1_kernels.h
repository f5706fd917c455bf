// bc1_kernels.h -- the BC1/BC3 encode kernels and their launches for one
// interpolator, included by the per-interpolator translation units
// (bc1_ideal.hip, bc1_ideal_round.hip, bc1_nvidia.hip, bc1_amd.hip), each of
// which instantiates them for its interpolator.
//
// Three encoders of BC1Encoder::EncodeBlock (quicktex/s3tc/bc1/BC1Encoder.cpp:261-342):
// k_bc1_encode (one block per 16 lanes, bc1_device.h), k_bc1_lane (one block per
// lane, bc1_lane.h) and k_bc1_lane_queue (one block per lane, blocks dealt from
// work queues, long endpoint searches parked for k_bc1_search_drain); the batch
// forms k_color_batch and k_lane_batch run the first two over a job list.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>

#include "bc1_device.h"
#include "bc1_lane.h"
#include "bc1_launch.h"
#include "kernels.h"
#include "kernels_dev.h"

namespace qtx {
namespace dev {

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

// One BC1 (or BC3) block per lane: encode and store.
template <int INTERP>
__device__ __forceinline__ void encode_lane_block(const uint8_t *__restrict__ rgba, int w, int h, int bw, int gid,
                                                  uint8_t *__restrict__ out, const DeviceTables *__restrict__ T,
                                                  const BC1Settings &s, bool bc3) {
    uint32_t px[16];
    load_block16(rgba, w, h, bw, gid, px);
    store_lane_block(out, (uint32_t)gid, px, lane::encode_bc1<INTERP, false>(px, s, T), bc3);
}

// Block lists of the static lane kernel.  A solid block (one RGBA value) costs
// EncodeBlock a few dozen instructions -- it is written from the single-colour
// tables (BC1Encoder.cpp:262-268) -- and any other block thousands; with one block
// per lane, each solid block would idle its lane while the wave encodes the rest.
// k_bc1_split stores the solid blocks and lists the indices of the others, and
// k_bc1_lane<..., LIST> encodes the lists, one listed block per lane, so that only
// a list's last wave runs short.  A split workgroup takes kSplitBlocks blocks and
// appends its listed ones to list (workgroup % kLaneQueues) with one atomic on that
// list's count (kLaneQueues counts, kLaneQueueStride words apart, in the launch's
// queue set: one address takes only so many atomics per second); list q holds
// list_capacity(nblocks) words from q x that capacity.  The encode's workgroup e
// takes entries of list e % kLaneQueues (workgroups go to the XCDs round-robin, so
// a list is split and encoded on one XCD).  kLaneRetireWord counts the encode
// workgroups that are done (each read its count first); the last of them returns
// the counts and itself to zero, so the next launch on the set (stream-ordered
// after this one) starts from zero.
template <int INTERP, bool BC3>
__global__ __launch_bounds__(256) void k_bc1_split(const uint8_t *__restrict__ rgba, int w, int h, int bw, int nblocks,
                                                   uint8_t *__restrict__ out, const DeviceTables *__restrict__ T,
                                                   uint32_t *__restrict__ queue, uint32_t *__restrict__ list) {
    __shared__ uint32_t staged[kSplitBlocks];
    __shared__ uint32_t nstaged, base;
    if (threadIdx.x == 0) nstaged = 0;
    __syncthreads();
    const uint32_t first = blockIdx.x * kSplitBlocks;
#pragma unroll 1
    for (uint32_t j = 0; j < kSplitBlocks; j += 256) {
        const uint32_t gid = first + j + threadIdx.x;
        bool listed = false;
        if (gid < (uint32_t)nblocks) {
            uint32_t px[16];
            load_block16(rgba, w, h, bw, (int)gid, px);
            bool solid = true;
#pragma unroll
            for (int i = 1; i < 16; i++) solid &= px[i] == px[0];  // alpha included (ColorBlock.h:80-86)
            if (solid) store_lane_block(out, gid, px, lane::write_solid_in(T->match[INTERP], px[0]), BC3);
            else listed = true;
        }
        const uint64_t mask = __ballot(listed);
        uint32_t at = 0;
        if ((threadIdx.x & 63) == 0 && mask) at = atomicAdd(&nstaged, (uint32_t)__popcll(mask));
        at = (uint32_t)__builtin_amdgcn_readfirstlane((int)at);
        if (listed) staged[at + __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u))] = gid;
    }
    __syncthreads();
    const uint32_t q = blockIdx.x % kLaneQueues;
    if (threadIdx.x == 0) base = nstaged ? atomicAdd(&queue[q * kLaneQueueStride], nstaged) : 0u;
    __syncthreads();
    uint32_t *const dst = list + (size_t)q * list_capacity((uint32_t)nblocks) + base;
    for (uint32_t i = threadIdx.x; i < nstaged; i += 256) dst[i] = staged[i];
}

// BC1/BC3 encode, one block per lane (any settings; LIGHT: the settings of
// levels 0-12 as constants, lane::light_settings).  LIST: the blocks k_bc1_split
// listed (the grid covers every list's capacity; workgroups past a list's end
// encode nothing).
template <int INTERP, bool BC3, bool LIGHT, bool LIST>
__global__ __launch_bounds__(256, 4) void k_bc1_lane(const uint8_t *__restrict__ rgba, int w, int h, int bw, int nblocks,
                                                  uint8_t *__restrict__ out, const DeviceTables *__restrict__ T,
                                                  BC1Settings s_in, uint32_t *__restrict__ queue,
                                                  const uint32_t *__restrict__ list) {
    const BC1Settings s = lane::fixed<LIGHT ? lane::kFixLight : lane::kFixNone>(s_in);
    if constexpr (LIST) {
        // the count, this lane's entry and the tables are read together; the entry
        // (inside the list's capacity, so always in bounds) is used only below the
        // count
        __shared__ uint32_t list_count;
        const uint32_t q = blockIdx.x % kLaneQueues;
        const uint32_t first = (blockIdx.x / kLaneQueues) * blockDim.x, i = first + threadIdx.x;
        if (threadIdx.x == 0)
            list_count = __hip_atomic_load(&queue[q * kLaneQueueStride], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const uint32_t gid = list[(size_t)q * list_capacity((uint32_t)nblocks) + i];
        lane::load_tables<INTERP, false>(T);
        __syncthreads();
        const uint32_t count = list_count;
        if (first < count && i < count) encode_lane_block<INTERP>(rgba, w, h, bw, (int)gid, out, T, s, BC3);
        // every workgroup counts itself once it has its count (thread 0 read it
        // above, before the barrier); the last returns the counts to zero
        if (threadIdx.x == 0 && atomicAdd(&queue[kLaneRetireWord], 1u) == gridDim.x - 1) {
            for (int k = 0; k < kLaneQueues; k++) atomicExch(&queue[k * kLaneQueueStride], 0u);
            atomicExch(&queue[kLaneRetireWord], 0u);
        }
    } else {
        lane::load_tables<INTERP, false>(T);
        __syncthreads();
        const int gid = (int)(blockIdx.x * blockDim.x + threadIdx.x);
        if (gid < nblocks) encode_lane_block<INTERP>(rgba, w, h, bw, gid, out, T, s, BC3);
    }
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

}  // namespace dev

namespace bc1k {
constexpr int kThreads = 256;
constexpr size_t kBc1Lds = sizeof(dev::Lds);
// the lane kernels' LDS is static (dev::lane::LaneLdsPal / LaneLdsPlain): no dynamic bytes
constexpr size_t kLaneLds = 0;

template <int I, bool B> void launch_queue(const Bc1Launch &a) {
    const dim3 grid(a.grid), block(a.threads);
    switch (a.fix) {
        case dev::lane::kFixHQ:
            hipLaunchKernelGGL((dev::k_bc1_lane_queue<I, B, dev::lane::kFixHQ>), grid, block, 0, a.stream, a.rgba, a.w,
                               a.h, a.bw, (uint32_t)a.nblocks, a.out, a.tables, a.s, a.queue, a.park, a.stop);
            break;
        case dev::lane::kFixLight:
            hipLaunchKernelGGL((dev::k_bc1_lane_queue<I, B, dev::lane::kFixLight>), grid, block, 0, a.stream, a.rgba,
                               a.w, a.h, a.bw, (uint32_t)a.nblocks, a.out, a.tables, a.s, a.queue, a.park, a.stop);
            break;
        default:
            hipLaunchKernelGGL((dev::k_bc1_lane_queue<I, B, dev::lane::kFixNone>), grid, block, 0, a.stream, a.rgba,
                               a.w, a.h, a.bw, (uint32_t)a.nblocks, a.out, a.tables, a.s, a.queue, a.park, a.stop);
            break;
    }
    if (a.drain)
        hipLaunchKernelGGL((dev::k_bc1_search_drain<I, B>), dim3(a.drain_grid), dim3(dev::lane::kQueueThreads), 0,
                           a.stream, a.rgba, a.w, a.h, a.bw, a.out, a.tables, a.s, a.park, a.queue);
}

template <int I, bool B, bool LIGHT> void launch_lane_build(const Bc1Launch &a) {
    if (a.list) {
        const uint32_t cap = list_capacity((uint32_t)a.nblocks);
        hipLaunchKernelGGL((dev::k_bc1_split<I, B>), dim3((a.nblocks + kSplitBlocks - 1) / kSplitBlocks), dim3(kThreads),
                           0, a.stream, a.rgba, a.w, a.h, a.bw, a.nblocks, a.out, a.tables, a.queue, a.list);
        hipLaunchKernelGGL((dev::k_bc1_lane<I, B, LIGHT, true>), dim3(kLaneQueues * (cap / kThreads)), dim3(kThreads),
                           kLaneLds, a.stream, a.rgba, a.w, a.h, a.bw, a.nblocks, a.out, a.tables, a.s, a.queue, a.list);
    } else {
        hipLaunchKernelGGL((dev::k_bc1_lane<I, B, LIGHT, false>), dim3(a.grid), dim3(kThreads), kLaneLds, a.stream,
                           a.rgba, a.w, a.h, a.bw, a.nblocks, a.out, a.tables, a.s, nullptr, nullptr);
    }
}

template <int I, bool B> void launch_lane(const Bc1Launch &a) {
    if (a.light) launch_lane_build<I, B, true>(a);
    else launch_lane_build<I, B, false>(a);
}

template <int I, bool B> void launch_group(const Bc1Launch &a) {
    hipLaunchKernelGGL((dev::k_bc1_encode<I, B>), dim3(a.grid), dim3(kThreads), kBc1Lds, a.stream, a.rgba, a.w, a.h,
                       a.bw, a.nblocks, a.out, a.tables, a.s);
}
}  // namespace bc1k

template <int I> hipError_t launch_bc1_interp(const Bc1Launch &a) {
    switch (a.route) {
        case kRouteQueue:
            if (a.bc3) bc1k::launch_queue<I, true>(a);
            else bc1k::launch_queue<I, false>(a);
            break;
        case kRouteLane:
            if (a.bc3) bc1k::launch_lane<I, true>(a);
            else bc1k::launch_lane<I, false>(a);
            break;
        default:
            if (a.bc3) bc1k::launch_group<I, true>(a);
            else bc1k::launch_group<I, false>(a);
            break;
    }
    return hipGetLastError();
}

template <int I>
hipError_t launch_bc1_batch_interp(bool lane, const BatchJob *d_jobs, int njobs, uint32_t nunits, uint32_t *d_queue,
                                   unsigned grid, const DeviceTables *d_tables, hipStream_t stream) {
    const dim3 block(bc1k::kThreads);
    if (lane)
        hipLaunchKernelGGL((dev::k_lane_batch<I>), dim3(std::max(1u, std::min(grid, nunits))), block, bc1k::kLaneLds,
                           stream, d_jobs, njobs, nunits, d_tables);
    else
        hipLaunchKernelGGL((dev::k_color_batch<I>), dim3(grid), block, bc1k::kBc1Lds, stream, d_jobs, njobs, nunits,
                           d_queue, d_tables);
    return hipGetLastError();
}

}  // namespace qtx
