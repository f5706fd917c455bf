// server.cpp -- the serving form of the codec (include/qtx.h, qtx_server_*).
//
// One loop thread per server.  Submitted requests wait in arrival order; when a
// batch slot is free the loop admits from the front of the queue every request
// whose device footprint (a copy of a pageable input) fits in the device arena
// and, for a destination in pageable memory, whose output fits in the pinned
// staging area, stopping at the first that does not fit so that no request is
// overtaken forever.  The admitted requests form
// one batch (at most batch_bytes of traffic, so that several batches fit).  A
// batch runs in order on one of two streams -- copies of pageable inputs and the
// job table, the light kernel (decodes, BC4/BC5), the BC1/BC3 kernels -- and
// slots alternate between the streams, so the work of consecutive batches
// overlaps and a small request admitted behind a long level-18 batch does not
// wait for it (the colour kernels of a served batch leave one workgroup slot per
// CU free for that).  The loop goes straight back to admitting: up to `slots`
// batches are on the device at once, and requests that arrive while a batch runs
// join the next one.  A request whose event has fired is complete and
// returns its memory; a slot is free again when all of its requests are.
//
// A request still waiting can be cancelled.  A batch that fails on an error the
// runtime raised before running its work fails only its own requests; an error
// from work on the device ends the server (srv::fault_scope).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "batch_internal.h"
#include "kernels.h"
#include "qtx_internal.h"
#include "runtime.h"
#include "server_core.h"

using namespace qtx;
using namespace qtx::rt;

namespace {

constexpr size_t kAlign = 256;
// resident workgroups per CU a served batch's colour kernels take (of the four
// that fit): the rest is room for the next batch's
constexpr int kServedWgPerCu = bt::kMaxColourWgPerCu - 1;
constexpr int kQueueWords = 2 * bt::kGroups;  // two words per kernel group
// pageable inputs from this size go through the loop's pinned chunks and copy
// threads (the runtime's own path for pageable memory moves about half as fast)
constexpr size_t kStageMin = (size_t)1 << 20;

size_t align_up(size_t n) { return (n + kAlign - 1) / kAlign * kAlign; }

struct Req {
    uint64_t ticket = 0;
    qtx_job job{};  // host pointers
    bool admitted = false;
    bt::Prepared prep{};
    bool dst_pinned = false;
    bool src_direct = false;  // the kernel reads the caller's pinned input in place
    bool light = false;       // served by the decode / BC4 / BC5 kernel, else by a BC1/BC3 kernel
    size_t dev_off = 0, dev_len = 0;  // device copy of a pageable input
    size_t stage_off = 0;
    bool staged = false;  // the kernel writes the output to pinned staging (pageable destination)
    bool done = false;
    qtx_server_result res{};
};
using ReqP = std::shared_ptr<Req>;

struct Slot {
    hipEvent_t ev_k = nullptr;   // its BC1/BC3 kernels are done: their outputs are in host memory
    hipEvent_t ev_l = nullptr;   // its decode / BC4 / BC5 kernel is done
    BatchJob *d_jobs = nullptr;
    BatchJob *h_jobs = nullptr;  // pinned staging of the job table
    uint32_t *d_queues = nullptr;
    std::vector<ReqP> reqs;
    bool pend_k = false, pend_l = false;  // kernels of the batch not yet seen complete
    bool busy = false;
};

}  // namespace

struct qtx_server {
    qtx_ctx *ctx = nullptr;
    qtx_server_config cfg{};
    int bc1_path = 0;
    uint8_t *d_arena = nullptr;
    uint8_t *h_stage = nullptr;
    srv::Arena arena, stage;
    std::vector<Slot> slots;
    // pinned chunks and copy threads for pageable inputs and outputs (the loop's)
    rt::Stager stage_io;
    // the two streams batches run on: slot i's copies and kernels on s_k[i % 2];
    // the context's own (its stream and copy stream), not made for the server
    hipStream_t s_k[2] = {nullptr, nullptr};
    // test hook: the batch with this number (1 = the first) fails before any of
    // its work is enqueued, as a call the runtime refuses would (0: never)
    uint64_t inject_fault_batch = 0;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<ReqP> waiting;
    std::unordered_map<uint64_t, ReqP> tickets;
    uint64_t next_ticket = 1;
    bool stopping = false;
    bool failed = false;
    std::string fail_msg;
    std::thread loop;
    std::chrono::steady_clock::time_point origin;
    qtx_server_stats stats{};
    double busy_since = -1.0;

    double now() const {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - origin).count();
    }
    void sync_streams() const {
        for (hipStream_t st : s_k)
            if (st) (void)hipStreamSynchronize(st);
    }
};

namespace {

// Device memory a request needs: a copy of its input when that is pageable (a
// pinned input is read in place; every output is written straight into host memory).
size_t footprint(const Req &r) { return r.src_direct ? 0 : align_up(r.prep.src_bytes); }

// Enqueue one batch (called without the lock; touches only the slot, the
// requests, which the loop owns while they are admitted, and the slot's stream,
// which only the loop uses).  Everything of slot i runs in order on s_k[i % 2]:
// the copies of pageable inputs and the job table, then the light kernel (decodes,
// BC4/BC5: short and PCIe-bound, so its requests complete first), then the
// BC1/BC3 kernels.  Kernels read pinned inputs in place over PCIe and store every
// output straight into host memory -- the caller's pinned buffer or the staging
// area -- so a batch costs one table upload and its kernels, whatever its size;
// the batches of the two streams overlap.
hipError_t enqueue(qtx_server *s, Slot &sl, int slot_index, uint64_t batch_no, const std::vector<ReqP> &reqs) {
    if (s->inject_fault_batch && batch_no == s->inject_fault_batch) return hipErrorInvalidValue;
    const hipStream_t st = s->s_k[slot_index & 1];
    std::vector<bt::Prepared> prep(reqs.size());
    hipError_t e = hipSuccess;
    for (size_t i = 0; i < reqs.size() && e == hipSuccess; i++) {
        const Req &r = *reqs[i];
        prep[i] = r.prep;
        if (r.src_direct) {
            prep[i].job.src = static_cast<const uint8_t *>(r.job.src);
        } else {
            uint8_t *dev = s->d_arena + r.dev_off;
            prep[i].job.src = dev;
            e = s->stage_io.upload(dev, static_cast<const uint8_t *>(r.job.src), r.prep.src_bytes, st, kStageMin);
        }
        prep[i].job.dst = r.staged ? s->h_stage + r.stage_off : static_cast<uint8_t *>(r.job.dst);
    }
    if (e != hipSuccess) return e;
    bt::Plan plan;
    const qtx_batch_tuning tuning{kServedWgPerCu, s->cfg.unit_kops};
    if (bt::plan(prep, tuning, s->ctx->cu_count, s->bc1_path, plan) != QTX_OK) return hipErrorInvalidValue;
    std::memcpy(sl.h_jobs, plan.host.data(), plan.host.size() * sizeof(BatchJob));
    e = hipMemcpyAsync(sl.d_jobs, sl.h_jobs, plan.host.size() * sizeof(BatchJob), hipMemcpyHostToDevice, st);
    sl.pend_k = sl.pend_l = false;
    for (int pass = 0; pass < 2 && e == hipSuccess; pass++) {  // the light group, then the colour groups
        for (size_t g = 0; g < plan.groups.size() && e == hipSuccess; g++) {
            bt::Plan::Group grp = plan.groups[g];
            const bool light = grp.kind == bt::kGroupLight;
            if (light != (pass == 0)) continue;
            (light ? sl.pend_l : sl.pend_k) = true;
            // PCIe-bound: leave the CUs to the encoders
            if (light) grp.grid = std::min<unsigned>(grp.grid, 2u * (unsigned)s->ctx->cu_count);
            // a colour batch keeps one workgroup slot per CU free for the batch on the
            // other stream, so a request admitted behind a long batch starts at once
            else grp.grid = std::min<unsigned>(grp.grid, (unsigned)(kServedWgPerCu * s->ctx->cu_count));
            e = bt::enqueue_group(grp, sl.d_jobs, sl.d_queues + 2 * g, s->ctx->d_tables, st);
        }
        if (e == hipSuccess && pass == 0 && sl.pend_l) e = hipEventRecord(sl.ev_l, st);
    }
    if (e == hipSuccess && sl.pend_k) e = hipEventRecord(sl.ev_k, st);
    return e;
}

void release(qtx_server *s, Req &r) {
    if (r.dev_len) s->arena.release(r.dev_off, r.dev_len);
    if (r.staged) s->stage.release(r.stage_off, r.prep.dst_bytes);
}

void finish(qtx_server *s, Req &r, qtx_status st, const char *msg, double t) {
    r.res.status = st;
    if (msg) std::snprintf(r.res.message, sizeof r.res.message, "%s", msg);
    r.res.done_s = t;
    r.done = true;
    if (st == QTX_OK) {
        s->stats.completed++;
        s->stats.pixels += (uint64_t)r.job.width * (uint64_t)r.job.height;
    } else {
        s->stats.failed++;
    }
}

void update_busy(qtx_server *s, double t) {
    const bool any = std::any_of(s->slots.begin(), s->slots.end(), [](const Slot &x) { return x.busy; });
    if (any && s->busy_since < 0) s->busy_since = t;
    if (!any && s->busy_since >= 0) {
        s->stats.busy_s += t - s->busy_since;
        s->busy_since = -1.0;
    }
}

void run(qtx_server *s) {
    (void)hipSetDevice(s->ctx->device);
    std::unique_lock<std::mutex> lk(s->mu);
    for (;;) {
        // 1. completed kernels: their requests' outputs are in host memory
        for (Slot &sl : s->slots) {
            if (!sl.busy) continue;
            bool progressed = false;
            for (int which = 0; which < 2; which++) {
                bool &pend = which ? sl.pend_l : sl.pend_k;
                if (!pend) continue;
                const hipError_t q = hipEventQuery(which ? sl.ev_l : sl.ev_k);
                if (q == hipErrorNotReady) continue;
                pend = false;
                progressed = true;
                std::string msg;
                if (q != hipSuccess) {
                    msg = std::string("batch failed: ") + hipGetErrorString(q);
                    if (srv::fault_scope((int)q) == srv::FaultScope::Server) {
                        s->failed = true;
                        s->fail_msg = msg;
                    }
                    lk.unlock();
                    s->sync_streams();
                    lk.lock();
                }
                for (auto &r : sl.reqs) {
                    if (r->light != (which == 1) || r->done) continue;
                    if (q == hipSuccess && r->staged) {
                        lk.unlock();
                        s->stage_io.host_copy(static_cast<uint8_t *>(r->job.dst), s->h_stage + r->stage_off,
                                              r->prep.dst_bytes);
                        lk.lock();
                    }
                    release(s, *r);
                    finish(s, *r, q == hipSuccess ? QTX_OK : QTX_EDEVICE, q == hipSuccess ? nullptr : msg.c_str(),
                           s->now());
                }
            }
            if (!sl.pend_k && !sl.pend_l) {
                sl.reqs.clear();
                sl.busy = false;
                update_busy(s, s->now());
            }
            if (progressed) s->cv_done.notify_all();
        }
        // 2. a new batch on a free slot
        // a free slot, preferring one whose stream has no batch in flight
        Slot *free_slot = nullptr;
        int stream_busy[2] = {0, 0};
        for (size_t i = 0; i < s->slots.size(); i++) stream_busy[i & 1] += s->slots[i].busy ? 1 : 0;
        for (size_t i = 0; i < s->slots.size(); i++) {
            Slot &sl = s->slots[i];
            if (sl.busy) continue;
            if (!free_slot || (stream_busy[i & 1] == 0 && stream_busy[(free_slot - s->slots.data()) & 1] > 0))
                free_slot = &sl;
        }
        if (free_slot && !s->waiting.empty() && !s->failed) {
            std::deque<ReqP> taken;
            size_t batch_host_bytes = 0;
            auto reserve = [&](ReqP &r) {
                size_t off = 0, soff = 0;
                const size_t need = footprint(*r);
                // a batch moves at most batch_bytes (a larger request goes alone), so
                // several batches are in flight and their transfers and kernels overlap
                const size_t moved = r->prep.src_bytes + r->prep.dst_bytes;
                if (batch_host_bytes > 0 && batch_host_bytes + moved > s->cfg.batch_bytes) return false;
                if (need && !s->arena.alloc(need, off)) return false;
                if (!r->dst_pinned) {
                    if (!s->stage.alloc(r->prep.dst_bytes, soff)) {
                        if (need) s->arena.release(off, need);
                        return false;
                    }
                    r->stage_off = soff;
                    r->staged = true;
                }
                r->dev_off = off;
                r->dev_len = need;
                batch_host_bytes += moved;
                return true;
            };
            srv::admit_fifo(s->waiting, (size_t)s->cfg.max_batch_jobs, reserve, taken);
            if (!taken.empty()) {
                std::vector<ReqP> reqs(taken.begin(), taken.end());
                const double t_admit = s->now();
                for (auto &r : reqs) {
                    r->admitted = true;
                    r->res.admit_s = t_admit;
                    r->res.batch_jobs = (int32_t)reqs.size();
                }
                s->stats.batches++;
                const uint64_t batch_no = s->stats.batches;
                const int slot_index = (int)(free_slot - s->slots.data());
                s->stats.max_batch_jobs = std::max<uint64_t>(s->stats.max_batch_jobs, reqs.size());
                s->stats.device_bytes_peak = std::max(s->stats.device_bytes_peak, s->arena.used());
                free_slot->busy = true;
                update_busy(s, t_admit);
                lk.unlock();
                const hipError_t e = enqueue(s, *free_slot, slot_index, batch_no, reqs);
                lk.lock();
                const double t_launch = s->now();
                if (e != hipSuccess) {
                    // nothing of this batch can be trusted: wait for what of it was
                    // enqueued before its memory is reused and its requests fail
                    lk.unlock();
                    s->sync_streams();
                    lk.lock();
                    const std::string msg = std::string("batch enqueue failed: ") + hipGetErrorString(e);
                    if (srv::fault_scope((int)e) == srv::FaultScope::Server) {
                        s->failed = true;
                        s->fail_msg = msg;
                    }
                    for (auto &r : reqs) {
                        release(s, *r);
                        finish(s, *r, QTX_EDEVICE, msg.c_str(), t_launch);
                    }
                    free_slot->busy = false;
                    free_slot->pend_k = free_slot->pend_l = false;
                    update_busy(s, t_launch);
                    s->cv_done.notify_all();
                } else {
                    for (auto &r : reqs) r->res.launch_s = t_launch;
                    free_slot->reqs = std::move(reqs);
                }
                continue;  // look for completions and another free slot at once
            }
        }
        const bool inflight = std::any_of(s->slots.begin(), s->slots.end(), [](const Slot &x) { return x.busy; });
        if (s->failed && !inflight) {
            // fail everything still waiting
            const double t = s->now();
            for (auto &r : s->waiting) finish(s, *r, QTX_EDEVICE, s->fail_msg.c_str(), t);
            s->waiting.clear();
            s->cv_done.notify_all();
        }
        if (s->stopping && s->waiting.empty() && !inflight) break;
        if (inflight)
            s->cv_work.wait_for(lk, std::chrono::microseconds(50));
        else
            s->cv_work.wait(lk, [&] { return s->stopping || (!s->waiting.empty() && !s->failed); });
    }
}

void destroy(qtx_server *s) {
    (void)hipSetDevice(s->ctx->device);
    s->sync_streams();
    for (Slot &sl : s->slots) {
        for (hipEvent_t ev : {sl.ev_k, sl.ev_l})
            if (ev) (void)hipEventDestroy(ev);
        if (sl.d_jobs) (void)hipFree(sl.d_jobs);
        if (sl.d_queues) (void)hipFree(sl.d_queues);
        if (sl.h_jobs) (void)hipHostFree(sl.h_jobs);
    }
    if (s->d_arena) (void)hipFree(s->d_arena);
    if (s->h_stage) (void)hipHostFree(s->h_stage);
    delete s;
}

}  // namespace

extern "C" {

qtx_status qtx_server_start(qtx_ctx *ctx, const qtx_server_config *config, qtx_server **out) {
    if (!out) return fail(QTX_EINVAL, "null output pointer");
    *out = nullptr;
    qtx_status st = check_ctx(ctx);
    if (st != QTX_OK) return st;
    qtx_server_config c{};
    if (config) c = *config;
    if (!c.device_bytes) c.device_bytes = (size_t)2 << 30;
    if (!c.staging_bytes) c.staging_bytes = (size_t)256 << 20;
    if (!c.slots) c.slots = 6;
    if (!c.max_batch_jobs) c.max_batch_jobs = 1024;
    if (!c.max_queue) c.max_queue = 65536;
    if (!c.unit_kops) c.unit_kops = 64;
    if (!c.batch_bytes) c.batch_bytes = std::min(c.device_bytes / (size_t)c.slots, (size_t)256 << 20);
    if (c.slots < 1 || c.slots > 8) return fail(QTX_EINVAL, "slots must be in 1..8");
    if (c.max_batch_jobs < 1 || c.max_batch_jobs > 65536) return fail(QTX_EINVAL, "max_batch_jobs must be in 1..65536");
    if (c.max_queue < 1) return fail(QTX_EINVAL, "max_queue must be positive");
    if (c.unit_kops < 16 || c.unit_kops > 16384) return fail(QTX_EINVAL, "unit_kops must be in 16..16384");
    if (c.device_bytes < (1u << 20)) return fail(QTX_EINVAL, "device_bytes must be at least 1 MiB");

    auto *s = new qtx_server;
    s->ctx = ctx;
    s->cfg = c;
    s->bc1_path = ctx->bc1_path;
    s->origin = std::chrono::steady_clock::now();
    if (const char *f = std::getenv("QTX_SERVER_INJECT_FAULT")) s->inject_fault_batch = std::strtoull(f, nullptr, 10);
    hipError_t e = hipMalloc(&s->d_arena, c.device_bytes);
    if (e == hipSuccess) e = hipHostMalloc((void **)&s->h_stage, c.staging_bytes, hipHostMallocDefault);
    // two streams, the context's: the runtime gives a process a few hardware queues
    // (four by default) and lets further streams share them, and work on a shared
    // queue runs in order -- two streams more than the context's could put the
    // server's pair on one queue, and a small request behind a long batch
    if (e == hipSuccess) {
        std::lock_guard<std::mutex> lk(ctx->mu);
        e = ensure_copy_stream(ctx);
    }
    s->s_k[0] = ctx->stream;
    s->s_k[1] = ctx->copy_stream;
    s->slots.resize((size_t)c.slots);
    for (Slot &sl : s->slots) {
        if (e == hipSuccess) e = hipEventCreateWithFlags(&sl.ev_k, hipEventDisableTiming);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&sl.ev_l, hipEventDisableTiming);
        if (e == hipSuccess) e = hipMalloc(&sl.d_jobs, (size_t)c.max_batch_jobs * sizeof(BatchJob));
        if (e == hipSuccess)
            e = hipHostMalloc((void **)&sl.h_jobs, (size_t)c.max_batch_jobs * sizeof(BatchJob), hipHostMallocDefault);
        if (e == hipSuccess) e = hipMalloc(&sl.d_queues, kQueueWords * sizeof(uint32_t));
        if (e == hipSuccess) e = hipMemset(sl.d_queues, 0, kQueueWords * sizeof(uint32_t));
    }
    if (e != hipSuccess) {
        const qtx_status r = e == hipErrorOutOfMemory ? fail(QTX_ENOMEM, "server set-up: %s", hipGetErrorString(e))
                                                      : hip_fail(e, "server set-up");
        destroy(s);
        return r;
    }
    s->arena.reset(c.device_bytes, kAlign);
    s->stage.reset(c.staging_bytes, kAlign);
    s->loop = std::thread(run, s);
    *out = s;
    return QTX_OK;
}

qtx_status qtx_server_submit(qtx_server *s, const qtx_job *job, size_t dst_bytes, uint64_t *ticket) {
    if (!s || !job || !ticket) return fail(QTX_EINVAL, "null argument");
    auto r = std::make_shared<Req>();
    r->job = *job;
    qtx_status st = bt::prepare(*job, bt::kNoIndex, r->prep);
    if (st != QTX_OK) return st;
    if (dst_bytes < r->prep.dst_bytes)
        return fail(QTX_EINVAL, "output buffer holds %zu bytes, %zu needed", dst_bytes, r->prep.dst_bytes);
    (void)hipSetDevice(s->ctx->device);
    r->dst_pinned = rt::host_pinned(job->dst);
    r->src_direct = rt::host_pinned(job->src);
    r->light = r->prep.group == bt::kGroupLight;
    const size_t need = footprint(*r);
    if (need > s->arena.capacity())
        return fail(QTX_ENOMEM, "request needs %zu bytes of device memory; the server has %zu", need,
                    s->arena.capacity());
    if (!r->dst_pinned && align_up(r->prep.dst_bytes) > s->stage.capacity())
        return fail(QTX_ENOMEM, "output of %zu bytes to pageable memory exceeds the %zu-byte staging area",
                    r->prep.dst_bytes, s->stage.capacity());
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->stopping) return fail(QTX_EINVAL, "server is stopping");
    if (s->failed) return fail(QTX_EDEVICE, "server failed: %s", s->fail_msg.c_str());
    if (s->waiting.size() >= (size_t)s->cfg.max_queue)
        return fail(QTX_EBUSY, "queue full (%d requests waiting)", s->cfg.max_queue);
    r->ticket = s->next_ticket++;
    r->res.submit_s = s->now();
    s->waiting.push_back(r);
    s->tickets[r->ticket] = r;
    s->stats.submitted++;
    s->stats.max_waiting = std::max<uint64_t>(s->stats.max_waiting, s->waiting.size());
    *ticket = r->ticket;
    s->cv_work.notify_one();
    return QTX_OK;
}

qtx_status qtx_server_wait(qtx_server *s, uint64_t ticket, int32_t timeout_ms, qtx_server_result *result) {
    if (!s || !result) return fail(QTX_EINVAL, "null argument");
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->tickets.find(ticket);
    if (it == s->tickets.end()) return fail(QTX_EINVAL, "unknown ticket %llu", (unsigned long long)ticket);
    ReqP r = it->second;
    if (timeout_ms < 0)
        s->cv_done.wait(lk, [&] { return r->done; });
    else
        s->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), [&] { return r->done; });
    if (!r->done) return fail(QTX_EBUSY, "request %llu not finished", (unsigned long long)ticket);
    *result = r->res;
    s->tickets.erase(ticket);
    return QTX_OK;
}

qtx_status qtx_server_cancel(qtx_server *s, uint64_t ticket) {
    if (!s) return fail(QTX_EINVAL, "null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->tickets.find(ticket);
    if (it == s->tickets.end()) return fail(QTX_EINVAL, "unknown ticket %llu", (unsigned long long)ticket);
    ReqP r = it->second;
    if (r->done) return fail(QTX_EINVAL, "request %llu has finished", (unsigned long long)ticket);
    if (r->admitted || !srv::remove_waiting(s->waiting, r))
        return fail(QTX_EBUSY, "request %llu is running; it completes normally", (unsigned long long)ticket);
    r->res.status = QTX_ECANCELED;
    std::snprintf(r->res.message, sizeof r->res.message, "cancelled before admission");
    r->res.done_s = s->now();
    r->done = true;
    s->stats.cancelled++;
    s->cv_done.notify_all();
    return QTX_OK;
}

qtx_status qtx_server_stats_get(qtx_server *s, qtx_server_stats *stats) {
    if (!s || !stats) return fail(QTX_EINVAL, "null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    *stats = s->stats;
    if (s->busy_since >= 0) stats->busy_s += s->now() - s->busy_since;
    return QTX_OK;
}

double qtx_server_now(qtx_server *s) { return s ? s->now() : 0.0; }

qtx_status qtx_server_stop(qtx_server *s) {
    if (!s) return QTX_OK;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stopping = true;
    }
    s->cv_work.notify_all();
    if (s->loop.joinable()) s->loop.join();
    destroy(s);
    return QTX_OK;
}

}  // extern "C"
