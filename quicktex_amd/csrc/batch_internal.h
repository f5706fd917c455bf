// batch_internal.h -- job validation and work-unit planning shared by the
// batched form (batch.cpp) and the serving loop (server.cpp).
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <vector>

#include "qtx_internal.h"

namespace qtx {
namespace bt {

constexpr int kMaxColourWgPerCu = 4;   // __launch_bounds__(256, 4) of the colour kernel
constexpr uint32_t kMaxRounds = 1024;  // a colour unit holds at most 4 * kMaxRounds blocks

// Kernel groups of a batch: BC1/BC3 jobs with endpoint search (the 16-lane
// persistent encoder) by interpolator, BC1/BC3 jobs without it (the lane-per-block
// encoder) by interpolator, and the light kernel.
constexpr int kGroupColour = 0, kGroupLane = 4, kGroupLight = 8, kGroups = 9;

// A validated job: its kernel descriptor, estimated cost and byte sizes.
struct Prepared {
    BatchJob job;
    double block_ops;  // estimated ALU ops per block
    int group;         // kernel group: kGroupColour + interp, kGroupLane + interp or kGroupLight
    size_t src_bytes, dst_bytes;
};

// Kernel-ready layout: descriptors in group order, one queue per group.
struct Plan {
    struct Group {
        int kind = kGroupLight;  // kGroupColour, kGroupLane or kGroupLight
        int interp = 0;
        std::vector<int> order;  // indices into the prepared list, descending total cost
        size_t first = 0;        // offset of the group's descriptors in host[]
        uint32_t nunits = 0;
        unsigned grid = 0;
    };
    std::vector<BatchJob> host;
    std::vector<Group> groups;  // colour groups by interpolator, then the light group
};

constexpr size_t kNoIndex = ~(size_t)0;
// Validates job i (messages start "job <i>: "; none for i == kNoIndex) and fills p.
qtx_status prepare(const qtx_job &j, size_t i, Prepared &p);
// bc1_path: QTX_BC1_AUTO / _LANE / _GROUP (the context's choice)
qtx_status plan(const std::vector<Prepared> &jobs, const qtx_batch_tuning &t, int cu_count, int bc1_path, Plan &out);
// Launches one group's kernel; d_jobs holds plan.host, d_queue the group's two
// queue words (zero before the first launch; the kernels leave them zero).
hipError_t enqueue_group(const Plan::Group &g, const BatchJob *d_jobs, uint32_t *d_queue,
                         const DeviceTables *d_tables, hipStream_t s);
// Enqueues jobs (device buffers) as one batch on s with qtx_batch_create's default
// tuning: the plan's group kernels one after the other over the context's d_once
// buffer -- no graph, streams or allocations per call, unlike a qtx_batch.  Under
// ctx->mu, with every run of the context on s.  host_table receives the job table
// the copy to the device reads: the caller keeps it until s is synchronised.
qtx_status run_once(qtx_ctx *ctx, const std::vector<qtx_job> &jobs, hipStream_t s,
                    std::vector<BatchJob> &host_table);

}  // namespace bt
}  // namespace qtx
