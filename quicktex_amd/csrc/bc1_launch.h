// bc1_launch.h -- the BC1/BC3 encode launches of one interpolator (host side).
//
// The encode kernels are instantiated per interpolator in translation units of
// their own (bc1_ideal.hip, bc1_ideal_round.hip, bc1_nvidia.hip, bc1_amd.hip),
// which compile in parallel; kernels.hip chooses the encoder and the grid and
// calls the interpolator's launch here.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "qtx_internal.h"

namespace qtx {

// Which BC1/BC3 encode kernel a launch runs.
enum Bc1Route {
    kRouteGroup = 0,  // k_bc1_encode: one block per group of 16 lanes
    kRouteLane = 1,   // k_bc1_lane: one block per lane, one fixed block each (with a list:
                      // k_bc1_split, then k_bc1_lane over the listed non-solid blocks)
    kRouteQueue = 2,  // k_bc1_lane_queue (+ k_bc1_search_drain): blocks dealt from queues
};

struct Bc1Launch {
    const uint8_t *rgba;
    int w, h, bw, nblocks;
    uint8_t *out;
    const DeviceTables *tables;
    BC1Settings s;
    bool bc3;
    int route;      // Bc1Route
    unsigned grid;  // workgroups
    int threads;    // kRouteQueue: threads per workgroup (768, or 256 for small launches)
    bool light;     // kRouteLane: the level 0-12 build
    int fix;        // kRouteQueue: the encoder build (dev::lane::Fixed)
    bool drain;     // kRouteQueue: the drain kernel follows (searches park at `stop`)
    uint32_t stop;
    unsigned drain_grid;
    uint32_t *queue;
    uint4 *park;
    uint32_t *list;  // kRouteLane: the block list (nblocks words), or nullptr
    hipStream_t stream;
};

// Launches a's kernels for interpolator I; returns hipGetLastError().
template <int I> hipError_t launch_bc1_interp(const Bc1Launch &a);

// The batch's BC1/BC3 kernels for interpolator I: k_lane_batch when lane (d_queue
// unused), else k_color_batch.
template <int I>
hipError_t launch_bc1_batch_interp(bool lane, const BatchJob *d_jobs, int njobs, uint32_t nunits, uint32_t *d_queue,
                                   unsigned grid, const DeviceTables *d_tables, hipStream_t stream);

}  // namespace qtx
