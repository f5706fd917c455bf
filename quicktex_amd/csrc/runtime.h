// runtime.h -- host-side internals shared by the C ABI translation units
// (capi.cpp, batch.cpp, server.cpp).
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <mutex>

#include "qtx_internal.h"
#include "staging.h"

// A device, its work stream, its copy of the constant tables and the staging
// buffers of the host-memory entry points.
struct qtx_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    qtx::DeviceTables *d_tables = nullptr;
    int cu_count = 0;
    int bc1_path = 0;  // QTX_BC1_AUTO / _LANE / _GROUP
    // staging for the host entry points
    uint8_t *d_in = nullptr;
    size_t in_cap = 0;
    uint8_t *d_out = nullptr;
    size_t out_cap = 0;
    uint32_t *d_flag = nullptr;  // result word of the alpha test
    qtx::rt::Stager stager;      // pinned staging of large host transfers (under mu)
    // Row bands of the large host encodes (capi.cpp): a stream for their transfers
    // and events for each band's upload and encode; made on the first such encode
    // (under mu).  The runtime spreads a process's streams over a few hardware
    // queues, so a context that never bands takes no queue for them.
    static constexpr int kMaxBands = 8;
    hipStream_t copy_stream = nullptr;
    hipEvent_t band_up[kMaxBands] = {}, band_done[kMaxBands] = {};
    // The batches the host entry points run once (batch.cpp, bt::run_once): the
    // queue words of their kernel groups (zero between runs: the kernels leave them
    // so), then the job table; grown on demand (under mu).
    uint8_t *d_once = nullptr;
    size_t once_cap = 0;
    std::mutex mu;
    // Work-queue words of the queued BC1/BC3 lane kernel: kQueueSlots sets, each
    // lent to one launch at a time.  A set is free again once the event recorded
    // after its last launch has completed (the kernel leaves its words zero), so
    // concurrent launches -- on any streams -- never share one.
    static constexpr int kQueueSlots = 64;
    uint32_t *d_queues = nullptr;
    hipEvent_t queue_done[kQueueSlots] = {};  // created on a slot's first use
    // The stream of each set's last launch (stream_id in capi.cpp): a launch on the
    // same stream may take the set before that launch completes.
    uint64_t queue_stream[kQueueSlots] = {};
    // One buffer per set, lent with its words: the queued kernel's parked endpoint
    // searches, or the static lane kernel's block lists; grown on demand
    // (qtx::bc1_park_bytes) with stream-ordered allocations.
    void *d_park[kQueueSlots] = {};
    size_t park_cap[kQueueSlots] = {};
    std::mutex queue_mu;
    // This device's copy of the library's static scratch array (capi.cpp), the
    // queue words and park records of qtx_bc1_encode_device_noalloc; looked up at
    // context creation.
    uint8_t *d_scratch = nullptr;
};

namespace qtx {
namespace rt {

// Makes ctx->copy_stream (and the band events) if it is not there yet; under ctx->mu,
// the device current.
hipError_t ensure_copy_stream(qtx_ctx *ctx);

// Thread-local message for qtx_last_error(); returns st.
qtx_status fail(qtx_status st, const char *fmt, ...) __attribute__((format(printf, 2, 3)));
qtx_status hip_fail(hipError_t e, const char *what);
const char *last_error();

qtx_status check_dims(int32_t w, int32_t h);
qtx_status check_ctx(qtx_ctx *ctx);
// qtx_bc1_params (clamped) -> kernel settings and interpolator
qtx_status to_settings(const qtx_bc1_params *p, BC1Settings &s, int &interp);
// grow a device buffer to at least need bytes (contents are not kept)
qtx_status ensure(uint8_t *&buf, size_t &cap, size_t need);
size_t block_bytes(int32_t w, int32_t h, int bytes_per_block);
// A set of the context's queue words lent to one launch on `stream`: a set
// whose last launch has completed, else one last lent to a launch on the same
// stream (stream order makes it free by the time this launch runs); `words` is
// nullptr when every set is in use by other streams (or the stream is capturing
// a graph) and the caller launches the queue-less kernel.  commit() marks the set
// busy until the work enqueued on the stream so far has completed.  Holds the
// context's queue lock from construction to commit() or destruction.
class QueueLease {
  public:
    QueueLease(qtx_ctx *ctx, hipStream_t stream, bool wanted);
    ~QueueLease();
    uint32_t *words = nullptr;
    // The set's park buffer, at least `bytes` long (grown if need be), or nullptr
    // when there is no set, bytes is 0 or the allocation fails (the launch then
    // parks nothing).
    void *park = nullptr;
    void reserve_park(size_t bytes);
    hipError_t commit();

  private:
    qtx_ctx *ctx_;
    hipStream_t stream_;
    uint64_t sid_ = 0;
    int slot_ = -1;
    bool locked_ = false;
};
size_t raw_bytes(int32_t w, int32_t h);

}  // namespace rt
}  // namespace qtx

#define QTX_HIP(call)                                                   \
    do {                                                                \
        hipError_t e_ = (call);                                         \
        if (e_ != hipSuccess) return ::qtx::rt::hip_fail(e_, #call);    \
    } while (0)
