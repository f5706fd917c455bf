// kernels.h -- launchers of the gfx950 kernels (host side).
#pragma once
#include <hip/hip_runtime.h>

#include "qtx_internal.h"

#include <cstdlib>
#include <cstring>

namespace qtx {

// Which BC1/BC3 encoder serves a setting: QTX_BC1_PATH=lane or =group forces one
// (both give identical bytes); unset or anything else picks per setting.
enum Bc1Path { BC1_AUTO = 0, BC1_LANE = 1, BC1_GROUP = 2 };
inline Bc1Path bc1_path_override() {
    static const Bc1Path p = [] {
        const char *e = std::getenv("QTX_BC1_PATH");
        if (e && !std::strcmp(e, "lane")) return BC1_LANE;
        if (e && !std::strcmp(e, "group")) return BC1_GROUP;
        return BC1_AUTO;
    }();
    return p;
}
bool use_lane_encoder(const BC1Settings &s);

hipError_t launch_bc1_encode(const uint8_t *d_rgba, int w, int h, uint8_t *d_out, const DeviceTables *d_tables,
                             const BC1Settings &s, int interp, bool bc3, hipStream_t stream);
hipError_t launch_bc4_encode(const uint8_t *d_rgba, int w, int h, uint8_t *d_out, int ch0, int ch1, bool bc5,
                             hipStream_t stream);
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
// 256 blocks per unit, one workgroup per unit.
constexpr uint32_t kLaneUnitBlocks = 256;
hipError_t launch_lane_batch(int interp, const BatchJob *d_jobs, int njobs, uint32_t nunits,
                             const DeviceTables *d_tables, hipStream_t stream);
hipError_t launch_light_batch(const BatchJob *d_jobs, int njobs, uint32_t nunits, hipStream_t stream);

}  // namespace qtx
