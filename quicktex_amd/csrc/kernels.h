// kernels.h -- launchers of the gfx950 kernels (host side).
#pragma once
#include <hip/hip_runtime.h>

#include "qtx_internal.h"

#include <cstdlib>
#include <cstring>

namespace qtx {

// Which BC1/BC3 encoder serves a setting (QTX_BC1_AUTO / _LANE / _GROUP): the
// context's choice, else the environment (QTX_BC1_PATH=lane or =group), else
// automatic.  Both encoders give identical bytes.
inline int bc1_path_env() {
    static const int p = [] {
        const char *e = std::getenv("QTX_BC1_PATH");
        if (e && !std::strcmp(e, "lane")) return (int)QTX_BC1_LANE;
        if (e && !std::strcmp(e, "group")) return (int)QTX_BC1_GROUP;
        return (int)QTX_BC1_AUTO;
    }();
    return p;
}
// QTX_LANE_QUEUE=0 in the environment: the lane encoder keeps one block per lane
// and takes no work queue (for comparisons; the bytes are the same).
inline bool lane_queue_env() {
    static const bool on = [] {
        const char *e = std::getenv("QTX_LANE_QUEUE");
        return !(e && !std::strcmp(e, "0"));
    }();
    return on;
}
// The static lane kernel takes its blocks from a list of the non-solid ones
// (k_bc1_split) at settings from kLaneListMinOps, unless QTX_LANE_LIST=0.
inline bool lane_list_env() {
    static const bool on = [] {
        const char *e = std::getenv("QTX_LANE_LIST");
        return !(e && !std::strcmp(e, "0"));
    }();
    return on;
}
// Launches of the queued kernel with at most a block per lane of 4-wave
// workgroups on every CU take such workgroups, unless QTX_QUEUE_SMALL=0.
inline bool queue_small_env() {
    static const bool on = [] {
        const char *e = std::getenv("QTX_QUEUE_SMALL");
        return !(e && !std::strcmp(e, "0"));
    }();
    return on;
}
// Estimated ALU ops per BC1 block for the encoder settings.
double bc1_block_ops(const BC1Settings &s);
// Whether the lane-per-block encoder serves settings s when `blocks` blocks of
// such work are launched together (else the 16-lane encoder does): the lane
// encoder needs enough blocks to fill the GPU, the 16-lane one spreads a block
// over 16 lanes and wins on few blocks of costly settings.
constexpr uint64_t kLaneFullBlocks = 262144;  // 256 CUs x 16 waves x 64 lanes
constexpr uint64_t kLaneMinBlocks = 32768;    // settings below kLaneCheapOps (levels 0-9)
constexpr double kLaneCheapOps = 10000.0;
constexpr double kLaneQueueBaseBlocks = 65536.0;  // costlier settings: from this + ops/2 blocks
bool use_lane_encoder(const BC1Settings &s, uint64_t blocks, int path);
// The same choice for a batch's BC1/BC3 jobs (blocks: the batch's total of such
// work), whose lane kernel is the static one.
constexpr double kLaneBatchMaxOps = 120000.0;
bool use_lane_batch(const BC1Settings &s, uint64_t blocks, int path);

// Work-queue words of the queued lane encoder: kLaneQueues block counters, each
// on its own 64-byte line, and a count of finished waves.
constexpr int kLaneQueues = 8;
constexpr int kLaneQueueStride = 16;
constexpr int kLaneRetireWord = kLaneQueues * kLaneQueueStride;
constexpr int kLaneQueueWords = 256;  // per queue, rounded up (1 KiB)
// The queued kernel serves settings from this estimated cost per block up; on
// cheaper settings the static kernel over its block lists is faster (4096x4096
// block mixture, FourColor: level 11, 16 k, 0.475 ms against 0.547 queued; level
// 12, 18 k, 0.525 against 0.585; level 13, 34 k, 0.542 against 0.533).
constexpr double kLaneQueueMinOps = 20000.0;
// The static kernel lists the non-solid blocks first from this estimated cost per
// block up; on cheaper settings the listing pass costs about what the idle lanes
// it saves do (same texture: level 7, 4 k, 0.166 ms listed against 0.177 plain;
// level 8, 6 k, 0.219 against 0.240; level 10, 12 k, 0.363 against 0.414; levels
// 2-6, 2.5 k, even; level 0, 2 k, 0.109 against 0.105).
constexpr double kLaneListMinOps = 3000.0;
// The encoder a launch of this texture takes (Bc1Route of bc1_launch.h; the
// queued route falls back to the static lane kernel when no queue set is lent).
int bc1_route(const BC1Settings &s, int w, int h, int path);
// The static lane kernel's block lists (k_bc1_split): a split workgroup takes
// kSplitBlocks blocks and appends the non-solid ones to one of kLaneQueues lists,
// each list_capacity(nblocks) words long (the blocks of every kLaneQueues-th
// split workgroup).
constexpr uint32_t kSplitBlocks = 1024;
__host__ __device__ constexpr uint32_t list_capacity(uint32_t nblocks) {
    return ((nblocks + kSplitBlocks - 1) / kSplitBlocks + kLaneQueues - 1) / kLaneQueues * kSplitBlocks;
}
// Whether launch_bc1_encode would use queue words for this texture -- the queued
// kernel's queues, or the counters of the static lane kernel's block list -- so
// that the caller need not lend a set otherwise.
bool wants_lane_queue(const BC1Settings &s, int w, int h, int path);
// Endpoint searches that go on past round kParkStop are parked by the queued
// kernel -- their state written to a park buffer, their lane given the next block
// -- and finished by a second kernel that runs one round per step on every lane
// and hands a lane the next parked search as soon as its own is over, so that a
// long search never holds 63 idle lanes.  Most searches end at round 33 or 34 (no
// improvement); the park words count the parked searches and the drain's
// finished waves.
constexpr uint32_t kParkStop = 34;
constexpr int kParkCountWord = 192;
constexpr int kParkRetireWord = 193;
constexpr size_t kParkRecordBytes = 32;
// kParkStop, or QTX_PARK_STOP when set (0: no parking)
inline uint32_t park_stop_env() {
    static const uint32_t stop = [] {
        const char *e = std::getenv("QTX_PARK_STOP");
        return e ? (uint32_t)std::strtoul(e, nullptr, 10) : kParkStop;
    }();
    return stop;
}
// Bytes of the buffer lent with the queue words that a launch of this texture
// needs (0: none): the queued kernel's park records (one per block at most, when
// its searches can outrun kParkStop), or the static lane kernel's block lists
// (kLaneQueues x list_capacity words).
size_t bc1_park_bytes(const BC1Settings &s, int w, int h, int path);
// d_queue: kLaneQueueWords zeroed words that no other launch uses until this one
// has completed (the lane encoder then deals blocks from queues, or lists the
// non-solid ones), or nullptr.  d_park: bc1_park_bytes bytes lent with d_queue
// (nullptr: searches run to the end where they started, and the static lane
// kernel takes every block).
hipError_t launch_bc1_encode(const uint8_t *d_rgba, int w, int h, uint8_t *d_out, const DeviceTables *d_tables,
                             const BC1Settings &s, int interp, bool bc3, int path, uint32_t *d_queue, void *d_park,
                             int cu_count, hipStream_t stream);
hipError_t launch_bc4_encode(const uint8_t *d_rgba, int w, int h, uint8_t *d_out, int ch0, int ch1, bool bc5,
                             hipStream_t stream);
// Zeroes n 32-bit words of device memory with a kernel on stream.
hipError_t launch_zero_words(uint32_t *d_words, uint32_t n, hipStream_t stream);
// format: QTX_FMT_*; p0/p1: write_alpha for BC1, channels for BC4/BC5
hipError_t launch_decode(int format, const uint8_t *d_blocks, int w, int h, uint8_t *d_rgba, int interp, int p0, int p1,
                         hipStream_t stream);

// Batch kernels (jobs sorted by descending cost, unit_begin ascending).
// Colour (BC1/BC3 encode) jobs of one interpolator: persistent grid taking work
// units from d_queue (two words, zero before the first launch; the kernel leaves
// them zero).  Light jobs (BC4/BC5 encode, all decoders): 256 blocks per unit,
// one workgroup per unit.
constexpr uint32_t kLightUnitBlocks = 256;
hipError_t launch_color_batch(int interp, const BatchJob *d_jobs, int njobs, uint32_t nunits, uint32_t *d_queue,
                              unsigned grid, const DeviceTables *d_tables, hipStream_t stream);
// BC1/BC3 jobs without endpoint search of one interpolator: one block per lane,
// 256 blocks per unit; grid workgroups (at most nunits) stride over the units.
constexpr uint32_t kLaneUnitBlocks = 256;
hipError_t launch_lane_batch(int interp, const BatchJob *d_jobs, int njobs, uint32_t nunits, unsigned grid,
                             const DeviceTables *d_tables, hipStream_t stream);
// grid: workgroups (at most nunits; they stride over the units)
hipError_t launch_light_batch(const BatchJob *d_jobs, int njobs, uint32_t nunits, unsigned grid, hipStream_t stream);

}  // namespace qtx
