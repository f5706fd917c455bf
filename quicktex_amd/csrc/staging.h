// staging.h -- host staging of the plain host entry points (qtx_bc1_encode, ...,
// qtx_bc5_decode).
//
// Their buffers are the caller's pageable memory.  The runtime moves pageable
// memory at about half the link's rate (it stages through buffers of its own:
// 28.5 GB/s measured for 64 MiB host to device, against 57 GB/s from pinned
// memory), so large transfers go through pinned chunks instead: a few host
// threads copy one chunk (72 GB/s with four) while the DMA engine moves the
// ones before it.
#pragma once
#include <hip/hip_runtime.h>

#include <condition_variable>
#include <cstddef>
#include <cstdint>
#include <mutex>
#include <thread>
#include <vector>

namespace qtx {
namespace rt {

// True when p lies in page-locked host memory the runtime knows (hipHostMalloc),
// which the DMA engine reads and writes in place.
bool host_pinned(const void *p);

// A piece of a device range and the host buffer it comes from or goes to: bytes
// [off, off + len) of the range.
struct HostPart {
    uint8_t *host;
    size_t off, len;
};

// memcpy split over the calling thread and `workers` threads kept for the purpose.
class CopyPool {
  public:
    explicit CopyPool(int workers);
    ~CopyPool();
    void copy(uint8_t *dst, const uint8_t *src, size_t n);

  private:
    void run(int part);
    std::mutex mu_;
    std::condition_variable go_, done_;
    std::vector<std::thread> threads_;
    uint8_t *dst_ = nullptr;
    const uint8_t *src_ = nullptr;
    size_t n_ = 0;
    uint64_t gen_ = 0;
    int pending_ = 0;
    bool stop_ = false;
};

// A context's staging: kSlots pinned chunks (64 MiB), their DMA completion events
// and the copy threads, made on first use.  One caller at a time (the context's
// host lock).  The threads fill slot after slot while the DMA engine drains them,
// and a slot is waited for only when it comes round again: an upload of up to 64
// MiB never waits on the GPU.  (Streams share the few hardware queues, so an event
// recorded on the copy stream can complete only after kernels of other streams:
// a wait for one in the middle of an upload could stall it behind an encode.)
// Host buffers that are pinned already go to the DMA engine directly.
class Stager {
  public:
    // 8 MiB chunks: measured on BC1 4096x4096 host encodes against 4 MiB ones, 4-6%
    // faster at levels 5-18 (fewer hand-offs between the copy threads and the DMA
    // engine); 16 MiB ones lengthen the pipeline's fill and drain.
    static constexpr int kSlots = 8;
    static constexpr size_t kChunk = (size_t)8 << 20;
    static constexpr size_t kMinBytes = (size_t)8 << 20;  // smaller uploads take the runtime's path
    // Downloads are staged from kMinDownBytes on, and one below four chunks moves in
    // four smaller ones (at least kMinChunk), so that the copy-out of one overlaps the
    // DMA of the next (decodes of 1448x1448: 0.63 -> 0.43 ms).  Uploads keep whole
    // chunks: there the copy in comes before the DMA, and smaller ones doubled the
    // time of a 4 MiB upload.
    static constexpr size_t kMinDownBytes = (size_t)2 << 20;
    static constexpr size_t kMinChunk = (size_t)512 << 10;
    static size_t chunk_for(size_t n) {
        if (n >= 4 * kChunk) return kChunk;
        const size_t q = ((n / 4) + 65535) & ~(size_t)65535;
        return q < kMinChunk ? kMinChunk : q;
    }
    static constexpr int kWorkers = 3;
    ~Stager();
    // Enqueues the copy of n bytes of host memory src to device memory d_dst on s,
    // staged from min_bytes on.  src is no longer read once it returns.
    hipError_t upload(uint8_t *d_dst, const uint8_t *src, size_t n, hipStream_t s, size_t min_bytes = kMinBytes);
    // Copies n bytes of device memory d_src to host memory dst after the work
    // already on s.  A staged copy (min_bytes and more, pageable dst) is complete
    // on return; any other is enqueued on s (dst holds the bytes once s is
    // synchronised).
    // min_bytes: the size from which the copy is staged (kMinBytes, or less where
    // the caller moves one transfer in several parts).
    hipError_t download(uint8_t *dst, const uint8_t *d_src, size_t n, hipStream_t s, size_t min_bytes = kMinDownBytes);
    // The same for a device range whose pieces come from (upload) or go to
    // (download) several host buffers: the range [0, n) moves in chunks and each
    // chunk's pieces are copied to or from their buffers, so many small buffers
    // cost one staged transfer.  Device bytes of the range that no part covers (the
    // callers' alignment gaps) receive unspecified values on the way up and are
    // not copied out on the way down.  Parts ascend by offset and do not overlap.  Below
    // kPartsMinBytes, or with every part pinned, each part moves on its own
    // (enqueued on s).
    static constexpr size_t kPartsMinBytes = (size_t)1 << 20;
    // memcpy between host buffers on the copy threads (large ones; small ones on the
    // caller's thread alone).
    void host_copy(uint8_t *dst, const uint8_t *src, size_t n);
    hipError_t upload_parts(uint8_t *d_dst, const HostPart *parts, size_t nparts, size_t n, hipStream_t s);
    hipError_t download_parts(const HostPart *parts, size_t nparts, const uint8_t *d_src, size_t n, hipStream_t s);
    // The first n bytes (at most kSlots * kChunk) of the pinned chunks, for a call
    // whose kernels read and write host memory in place; returns once the chunks'
    // earlier transfers are done.  They are the caller's until its next call here.
    hipError_t direct(size_t n, uint8_t **p);

  private:
    hipError_t init();
    uint8_t *pin_ = nullptr;
    hipEvent_t ev_[kSlots] = {};
    unsigned next_ = 0;  // the slot the next chunk takes, kept across calls
    CopyPool *pool_ = nullptr;
};

}  // namespace rt
}  // namespace qtx
