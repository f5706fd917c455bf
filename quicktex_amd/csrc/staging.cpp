// staging.cpp -- the pinned, chunked host staging of staging.h.
#include "staging.h"

#include <algorithm>
#include <cstring>

namespace qtx {
namespace rt {

bool host_pinned(const void *p) {
    hipPointerAttribute_t a{};
    if (hipPointerGetAttributes(&a, p) != hipSuccess) {
        (void)hipGetLastError();  // pageable memory reports an error; clear it
        return false;
    }
    return a.type == hipMemoryTypeHost;
}

CopyPool::CopyPool(int workers) {
    for (int i = 0; i < workers; i++) threads_.emplace_back([this, i] { run(i + 1); });
}

CopyPool::~CopyPool() {
    {
        std::lock_guard<std::mutex> lk(mu_);
        stop_ = true;
    }
    go_.notify_all();
    for (std::thread &t : threads_) t.join();
}

namespace {
// Part `part` of `parts` of [0, n), page-aligned boundaries.
void part_range(size_t n, int parts, int part, size_t &b, size_t &e) {
    size_t per = (n + (size_t)parts - 1) / (size_t)parts;
    per = (per + 4095) & ~(size_t)4095;
    b = std::min(n, per * (size_t)part);
    e = std::min(n, b + per);
}
}  // namespace

void CopyPool::run(int part) {
    uint64_t seen = 0;
    for (;;) {
        uint8_t *dst;
        const uint8_t *src;
        size_t n;
        {
            std::unique_lock<std::mutex> lk(mu_);
            go_.wait(lk, [&] { return stop_ || gen_ != seen; });
            if (stop_) return;
            seen = gen_;
            dst = dst_;
            src = src_;
            n = n_;
        }
        size_t b, e;
        part_range(n, (int)threads_.size() + 1, part, b, e);
        if (e > b) std::memcpy(dst + b, src + b, e - b);
        {
            std::lock_guard<std::mutex> lk(mu_);
            if (--pending_ == 0) done_.notify_one();
        }
    }
}

void CopyPool::copy(uint8_t *dst, const uint8_t *src, size_t n) {
    {
        std::lock_guard<std::mutex> lk(mu_);
        dst_ = dst;
        src_ = src;
        n_ = n;
        pending_ = (int)threads_.size();
        gen_++;
    }
    go_.notify_all();
    size_t b, e;
    part_range(n, (int)threads_.size() + 1, 0, b, e);
    if (e > b) std::memcpy(dst + b, src + b, e - b);
    std::unique_lock<std::mutex> lk(mu_);
    done_.wait(lk, [&] { return pending_ == 0; });
}

Stager::~Stager() {
    for (hipEvent_t &ev : ev_)
        if (ev) {
            (void)hipEventSynchronize(ev);
            (void)hipEventDestroy(ev);
        }
    if (pin_) (void)hipHostFree(pin_);
    delete pool_;
}

hipError_t Stager::init() {
    if (pin_) return hipSuccess;
    hipError_t e = hipHostMalloc((void **)&pin_, kSlots * kChunk, hipHostMallocDefault);
    if (e != hipSuccess) {
        pin_ = nullptr;
        return e;
    }
    for (hipEvent_t &ev : ev_) {
        if ((e = hipEventCreateWithFlags(&ev, hipEventDisableTiming)) != hipSuccess) {
            ev = nullptr;
            for (hipEvent_t &made : ev_)
                if (made) (void)hipEventDestroy(made), made = nullptr;
            (void)hipHostFree(pin_);
            pin_ = nullptr;
            return e;
        }
    }
    pool_ = new CopyPool(kWorkers);
    return hipSuccess;
}

void Stager::host_copy(uint8_t *dst, const uint8_t *src, size_t n) {
    if (n >= ((size_t)256 << 10) && init() == hipSuccess) pool_->copy(dst, src, n);
    else std::memcpy(dst, src, n);
}

hipError_t Stager::upload(uint8_t *d_dst, const uint8_t *src, size_t n, hipStream_t s, size_t min_bytes) {
    if (n < min_bytes || host_pinned(src) || init() != hipSuccess)
        return hipMemcpyAsync(d_dst, src, n, hipMemcpyHostToDevice, s);
    for (size_t off = 0; off < n; off += kChunk) {
        const size_t len = std::min(kChunk, n - off);
        const unsigned k = next_++ % kSlots;
        uint8_t *slot = pin_ + k * kChunk;
        // the slot's last DMA (kSlots chunks back, in this call or an earlier one) is done
        hipError_t e = hipEventSynchronize(ev_[k]);
        if (e != hipSuccess) return e;
        pool_->copy(slot, src + off, len);
        if ((e = hipMemcpyAsync(d_dst + off, slot, len, hipMemcpyHostToDevice, s)) != hipSuccess) return e;
        if ((e = hipEventRecord(ev_[k], s)) != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t Stager::download(uint8_t *dst, const uint8_t *d_src, size_t n, hipStream_t s, size_t min_bytes) {
    hipError_t e;
    if (n < min_bytes || host_pinned(dst) || init() != hipSuccess)
        return hipMemcpyAsync(dst, d_src, n, hipMemcpyDeviceToHost, s);
    const size_t C = chunk_for(n);
    const size_t chunks = (n + C - 1) / C;
    const unsigned first = next_;
    next_ += (unsigned)chunks;
    // chunk k goes to slot (first + k) % kSlots; the DMAs of the next kSlots - 1
    // chunks run while chunk k is copied out
    auto slot = [&](size_t k) { return (unsigned)(first + k) % kSlots; };
    auto enqueue = [&](size_t k) -> hipError_t {
        const size_t off = k * C, len = std::min(C, n - off);
        // the slot's last chunk (an upload's DMA, or this call's chunk k - kSlots,
        // copied out already) is done with it
        hipError_t r = hipEventSynchronize(ev_[slot(k)]);
        if (r == hipSuccess) r = hipMemcpyAsync(pin_ + slot(k) * kChunk, d_src + off, len, hipMemcpyDeviceToHost, s);
        if (r == hipSuccess) r = hipEventRecord(ev_[slot(k)], s);
        return r;
    };
    for (size_t k = 0; k < chunks && k + 1 < (size_t)kSlots; k++)
        if ((e = enqueue(k)) != hipSuccess) return e;
    for (size_t k = 0; k < chunks; k++) {
        if ((e = hipEventSynchronize(ev_[slot(k)])) != hipSuccess) return e;
        if (k + kSlots - 1 < chunks && (e = enqueue(k + kSlots - 1)) != hipSuccess) return e;
        const size_t off = k * C, len = std::min(C, n - off);
        pool_->copy(dst + off, pin_ + slot(k) * kChunk, len);
    }
    return hipSuccess;
}

hipError_t Stager::direct(size_t n, uint8_t **p) {
    if (n > kSlots * kChunk) return hipErrorInvalidValue;
    hipError_t e = init();
    if (e != hipSuccess) return e;
    for (size_t k = 0; k * kChunk < n; k++)
        if ((e = hipEventSynchronize(ev_[k])) != hipSuccess) return e;
    *p = pin_;
    return hipSuccess;
}

namespace {
// Copies between a chunk [c0, c0 + len) of the range and the parts it overlaps:
// to the parts (out) or from them into the chunk.
void copy_pieces(CopyPool *pool, uint8_t *chunk, size_t c0, size_t len, const HostPart *parts, size_t nparts,
                 size_t &first, bool out) {
    while (first < nparts && parts[first].off + parts[first].len <= c0) first++;
    for (size_t i = first; i < nparts && parts[i].off < c0 + len; i++) {
        const size_t b = std::max(parts[i].off, c0), e = std::min(parts[i].off + parts[i].len, c0 + len);
        if (e <= b) continue;
        uint8_t *host = parts[i].host + (b - parts[i].off), *slot = chunk + (b - c0);
        uint8_t *dst = out ? host : slot;
        const uint8_t *src = out ? slot : host;
        if (e - b >= ((size_t)256 << 10)) pool->copy(dst, src, e - b);
        else std::memcpy(dst, src, e - b);
    }
}

bool all_pinned(const HostPart *parts, size_t nparts) {
    for (size_t i = 0; i < nparts; i++)
        if (!host_pinned(parts[i].host)) return false;
    return true;
}
}  // namespace

hipError_t Stager::upload_parts(uint8_t *d_dst, const HostPart *parts, size_t nparts, size_t n, hipStream_t s) {
    if (n < kPartsMinBytes || all_pinned(parts, nparts) || init() != hipSuccess) {
        for (size_t i = 0; i < nparts; i++) {
            const hipError_t e = upload(d_dst + parts[i].off, parts[i].host, parts[i].len, s);
            if (e != hipSuccess) return e;
        }
        return hipSuccess;
    }
    size_t first = 0;
    for (size_t off = 0; off < n; off += kChunk) {
        const size_t len = std::min(kChunk, n - off);
        const unsigned k = next_++ % kSlots;
        uint8_t *slot = pin_ + k * kChunk;
        hipError_t e = hipEventSynchronize(ev_[k]);
        if (e != hipSuccess) return e;
        copy_pieces(pool_, slot, off, len, parts, nparts, first, false);
        if ((e = hipMemcpyAsync(d_dst + off, slot, len, hipMemcpyHostToDevice, s)) != hipSuccess) return e;
        if ((e = hipEventRecord(ev_[k], s)) != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t Stager::download_parts(const HostPart *parts, size_t nparts, const uint8_t *d_src, size_t n, hipStream_t s) {
    if (n < kPartsMinBytes || all_pinned(parts, nparts) || init() != hipSuccess) {
        for (size_t i = 0; i < nparts; i++) {
            const hipError_t e = download(parts[i].host, d_src + parts[i].off, parts[i].len, s);
            if (e != hipSuccess) return e;
        }
        return hipSuccess;
    }
    hipError_t e;
    const size_t chunks = (n + kChunk - 1) / kChunk;
    const unsigned first_slot = next_;
    next_ += (unsigned)chunks;
    auto slot = [&](size_t k) { return (unsigned)(first_slot + k) % kSlots; };
    auto enqueue = [&](size_t k) -> hipError_t {
        const size_t off = k * kChunk, len = std::min(kChunk, n - off);
        hipError_t r = hipEventSynchronize(ev_[slot(k)]);
        if (r == hipSuccess) r = hipMemcpyAsync(pin_ + slot(k) * kChunk, d_src + off, len, hipMemcpyDeviceToHost, s);
        if (r == hipSuccess) r = hipEventRecord(ev_[slot(k)], s);
        return r;
    };
    for (size_t k = 0; k < chunks && k + 1 < (size_t)kSlots; k++)
        if ((e = enqueue(k)) != hipSuccess) return e;
    size_t first = 0;
    for (size_t k = 0; k < chunks; k++) {
        if ((e = hipEventSynchronize(ev_[slot(k)])) != hipSuccess) return e;
        if (k + kSlots - 1 < chunks && (e = enqueue(k + kSlots - 1)) != hipSuccess) return e;
        const size_t off = k * kChunk, len = std::min(kChunk, n - off);
        copy_pieces(pool_, pin_ + slot(k) * kChunk, off, len, parts, nparts, first, true);
    }
    return hipSuccess;
}

}  // namespace rt
}  // namespace qtx
