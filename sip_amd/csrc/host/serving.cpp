// Serving form: a continuous-batching request loop for 1D FFT requests on one GPU.
//
// Requests (C2C or R2C, length n, batch b, device input; output either caller-provided or taken
// from the server's arena) enter a FIFO. A scheduler thread:
//   * retires finished launches (event query), records per-request latency (arrival -> observed
//     completion) and returns their memory to the arena;
//   * admits requests from the head of the queue while the HBM budget allows (outputs it owns,
//     workspace and row tables are carved from one arena allocated once);
//   * coalesces queued requests of the same (type, n) into one launch: all their rows go through a
//     single gathered-row plan (the row-offset tables are staged in pinned memory and copied
//     asynchronously), so the kernels are the ones a request would use alone and each result is
//     identical alone and in the mix;
//   * optionally (SIP_FFT_SERVE_MIXED=1) coalesces queued C2C requests of different short lengths
//     (2,3,5-smooth up to 4096) into one persistent multi-length launch (kernels/persist.hip; its
//     descriptors staged the same way); a short C2C request alone then runs the same launch with
//     one group, so here too each result is identical alone and in the mix;
//   * spreads launches over several streams (default priority), so work admitted while other work
//     runs overlaps on the device.
// Plans are rebuilt per launch on the host (microseconds); the coefficient tables of every shape
// served stay resident on the device for the server's lifetime (keep_resident).
// Extension API sipFftServe*; no counterpart in the reference.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <random>
#include <sstream>
#include <thread>
#include <unordered_map>

#include "persist.h"
#include "plan.h"

namespace sip {
namespace {

using clk = std::chrono::steady_clock;

double now_s()
{
    return std::chrono::duration<double>(clk::now().time_since_epoch()).count();
}

size_t granule(size_t bytes) { return bytes == 0 ? 256 : (bytes + 255) / 256 * 256; }

// First-fit allocator over one device arena (offsets; 256-byte granules).
class Arena {
public:
    explicit Arena(size_t bytes) : size_(bytes) { free_[0] = bytes; }
    // returns offset or SIZE_MAX
    size_t alloc(size_t bytes)
    {
        bytes = (bytes + 255) / 256 * 256;
        if (bytes == 0) bytes = 256;
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second >= bytes) {
                const size_t off = it->first, sz = it->second;
                free_.erase(it);
                if (sz > bytes) free_[off + bytes] = sz - bytes;
                used_ += bytes;
                return off;
            }
        }
        return SIZE_MAX;
    }
    void release(size_t off, size_t bytes)
    {
        bytes = (bytes + 255) / 256 * 256;
        if (bytes == 0) bytes = 256;
        used_ -= bytes;
        auto it = free_.emplace(off, bytes).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }
    size_t used() const { return used_; }
    size_t size() const { return size_; }

private:
    size_t size_;
    size_t used_ = 0;
    std::map<size_t, size_t> free_;
};

struct Req {
    uint64_t id = 0;
    int type = kTypeC2C;
    int64_t n = 0, batch = 0;
    const void *in = nullptr;
    void *out = nullptr;
    bool own_out = false;
    size_t out_off = SIZE_MAX, out_bytes = 0;
    size_t ws_need = 0;  // workspace of the request's plan when it runs alone
    double t_arr = 0, t_adm = 0, t_done = 0;
    bool done = false;
    bool detached = false;  // nobody waits for it: released when it completes (load generation)
    int32_t status = 0;
};

struct Launch {
    std::vector<Req *> reqs;
    int stream = 0;
    hipEvent_t ev = nullptr;
    size_t ws_off = SIZE_MAX, ws_bytes = 0;
    size_t tab_off = SIZE_MAX, tab_bytes = 0;
    size_t pin_off = 0, pin_bytes = 0;
    std::unique_ptr<Plan> plan;
    std::vector<TablePtr> hold;  // coefficient tables of a persistent launch
};

class Server {
public:
    Server(size_t budget, int nstreams, int64_t max_rows) : budget_(budget), nstreams_(nstreams), max_rows_(max_rows)
    {
        SIP_HIP_CHECK(hipGetDevice(&device_));
        SIP_HIP_CHECK(hipMalloc(&arena_base_, budget));
        arena_ = std::make_unique<Arena>(budget);
        pin_bytes_ = 64ull << 20;
        SIP_HIP_CHECK(hipHostMalloc(&pin_, pin_bytes_, hipHostMallocDefault));
        for (int i = 0; i < nstreams_; i++) {
            hipStream_t s;
            SIP_HIP_CHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
            streams_.push_back(s);
        }
        inflight_per_stream_.assign((size_t)nstreams_, 0);
        // resident coefficient tables live outside the arena: at most a sixteenth of the budget
        // (at least 64 MiB), or SIP_FFT_SERVE_RESIDENT_MB
        resident_cap_ = std::max<size_t>(size_t(64) << 20, budget / 16);
        if (const char *e = std::getenv("SIP_FFT_SERVE_RESIDENT_MB")) resident_cap_ = (size_t)std::atoll(e) << 20;
        lat_ms_.reserve(size_t(1) << 22);  // a 60 s stream at 50,000 req/s without regrowing under the lock
        worker_ = std::thread([this] { loop(); });
    }

    ~Server()
    {
        {
            std::lock_guard<std::mutex> lk(mu_);
            stop_ = true;
        }
        cv_.notify_all();
        if (worker_.joinable()) worker_.join();
        for (auto &l : inflight_) {
            (void)hipEventSynchronize(l->ev);
            (void)hipEventDestroy(l->ev);
        }
        for (hipEvent_t e : ev_pool_) (void)hipEventDestroy(e);
        for (hipStream_t s : streams_) (void)hipStreamDestroy(s);
        if (pin_) (void)hipHostFree(pin_);
        if (arena_base_) (void)hipFree(arena_base_);
    }

    // detached: the request is released when it completes (its latency still counts in the stats);
    // nothing may wait for it. Load generators use this so that millions of finished requests do
    // not accumulate in the ticket map.
    int32_t submit(int type, int64_t n, int64_t batch, const void *in, void *out, uint64_t *ticket,
                   bool detached = false)
    {
        std::unique_ptr<Req> r;
        const int32_t st = prepare(type, n, batch, in, out, r);
        if (st != AsdSip::ACL_SUCCESS) return st;
        r->detached = detached;
        r->t_arr = now_s();
        std::lock_guard<std::mutex> lk(mu_);
        enqueue_locked(std::move(r), ticket);
        cv_.notify_all();
        return AsdSip::ACL_SUCCESS;
    }

    // Several requests handed over together: every one is validated first (nothing is queued when one
    // is refused; `bad` receives its index), then all enter the queue under one lock, so the
    // scheduler sees them at once and can coalesce those of one shape.
    int32_t submit_many(int64_t count, const int32_t *types, const int64_t *ns, const int64_t *batches,
                        const void *const *ins, void *const *outs, uint64_t *tickets, int64_t *bad)
    {
        if (count <= 0 || !types || !ns || !batches || !ins) return AsdSip::ACL_ERROR_INVALID_PARAM;
        std::vector<std::unique_ptr<Req>> rs((size_t)count);
        for (int64_t i = 0; i < count; i++) {
            const int32_t st = prepare(types[i], ns[i], batches[i], ins[i], outs ? outs[i] : nullptr, rs[(size_t)i]);
            if (st != AsdSip::ACL_SUCCESS) {
                if (bad) *bad = i;
                return st;
            }
        }
        const double t = now_s();
        std::lock_guard<std::mutex> lk(mu_);
        for (int64_t i = 0; i < count; i++) {
            rs[(size_t)i]->t_arr = t;
            enqueue_locked(std::move(rs[(size_t)i]), tickets ? &tickets[i] : nullptr);
        }
        cv_.notify_all();
        return AsdSip::ACL_SUCCESS;
    }

    int32_t wait(uint64_t ticket, double *latency_ms)
    {
        std::unique_lock<std::mutex> lk(mu_);
        auto it = reqs_.find(ticket);
        if (it == reqs_.end()) return AsdSip::ACL_ERROR_INVALID_PARAM;
        Req *r = it->second.get();
        done_cv_.wait(lk, [&] { return r->done; });
        if (latency_ms) *latency_ms = (r->t_done - r->t_arr) * 1e3;
        const int32_t st = r->status;
        reqs_.erase(it);
        return st;
    }

    // Drops requests that are still queued (not admitted); they complete with status 200001.
    size_t cancel_queued()
    {
        std::lock_guard<std::mutex> lk(mu_);
        const size_t n = queue_.size();
        const double t = now_s();
        for (Req *r : queue_) {
            r->done = true;
            r->status = AsdSip::ACL_ERROR_API_NOT_SUPPORT;
            r->t_done = t;
            if (r->detached) reqs_.erase(r->id);
        }
        queue_.clear();
        done_cv_.notify_all();
        return n;
    }

    double last_done() const { return last_done_; }

    // Nothing queued and nothing running.
    bool idle()
    {
        std::lock_guard<std::mutex> lk(mu_);
        return queue_.empty() && inflight_.empty();
    }

    // Requests completed since creation: successful, failed.
    std::pair<uint64_t, uint64_t> completed()
    {
        std::lock_guard<std::mutex> lk(mu_);
        return {completed_ok_, completed_err_};
    }

    void drain()
    {
        std::unique_lock<std::mutex> lk(mu_);
        done_cv_.wait(lk, [&] { return queue_.empty() && inflight_.empty(); });
    }

    // latencies (ms) of requests completed since the last reset, and counters; the reset happens
    // only when the report fits `cap` bytes (a size query does not lose the data)
    std::string stats(bool reset, size_t cap = SIZE_MAX)
    {
        std::lock_guard<std::mutex> lk(mu_);
        std::vector<double> lat = lat_ms_;
        auto pct = [&](double q) {
            if (lat.empty()) return 0.0;
            const size_t k = (size_t)std::min<double>(lat.size() - 1, q * lat.size());
            std::nth_element(lat.begin(), lat.begin() + (std::ptrdiff_t)k, lat.end());
            return lat[k];
        };
        std::ostringstream os;
        os << "{\"completed\":" << lat.size() << ",\"p50_ms\":" << pct(0.50) << ",\"p90_ms\":" << pct(0.90)
           << ",\"p99_ms\":" << pct(0.99) << ",\"max_ms\":" << (lat.empty() ? 0.0 : *std::max_element(lat.begin(), lat.end()))
           << ",\"launches\":" << launches_ << ",\"mixed_launches\":" << mixed_launches_
           << ",\"coalesced_requests\":" << coalesced_ << ",\"queued\":" << queue_.size()
           << ",\"inflight_launches\":" << inflight_.size() << ",\"arena_used\":" << arena_->used()
           << ",\"arena_peak\":" << arena_peak_ << ",\"budget\":" << budget_ << ",\"bytes_done\":" << bytes_done_
           << ",\"flops_done\":" << flops_done_ << ",\"resident_tables\":" << resident_.size()
           << ",\"resident_bytes\":" << resident_bytes_ << ",\"resident_cap\":" << resident_cap_
           << ",\"resident_evictions\":" << resident_evictions_ << "}";
        if (reset && os.str().size() + 1 <= cap) {
            lat_ms_.clear();
            launches_ = coalesced_ = mixed_launches_ = 0;
            bytes_done_ = flops_done_ = 0;
            arena_peak_ = arena_->used();
        }
        return os.str();
    }

    int device() const { return device_; }

private:
    // A validated request, not yet queued. A request that could never be admitted alone is refused:
    // run alone, its launch needs its own output (when the server owns it) and the workspace of its
    // plain plan; the layout is host-only work of a few microseconds.
    int32_t prepare(int type, int64_t n, int64_t batch, const void *in, void *out, std::unique_ptr<Req> &out_req)
    {
        if ((type != kTypeC2C && type != kTypeR2C) || n <= 0 || n > kMaxFftSize || batch <= 0 || !in)
            return AsdSip::ACL_ERROR_INVALID_PARAM;
        auto r = std::make_unique<Req>();
        r->type = type;
        r->n = n;
        r->batch = batch;
        r->in = in;
        r->out = out;
        r->own_out = out == nullptr;
        r->out_bytes = (size_t)batch * (size_t)(type == kTypeC2C ? n : n / 2 + 1) * 8;
        try {
            Plan dry;
            dry.fftType = type;
            dry.direction = kDirForward;
            dry.big_leaf = false;  // as the launches below
            dry.sizes = {n};
            dry.batch = batch;
            plan_build(dry);
            r->ws_need = dry.ws_bytes;
        } catch (const SipError &e) {
            return e.status;
        }
        const size_t need = (r->own_out ? granule(r->out_bytes) : 0) + (r->ws_need ? granule(r->ws_need) : 0);
        if (need > budget_) return AsdSip::ACL_ERROR_INVALID_PARAM;
        out_req = std::move(r);
        return AsdSip::ACL_SUCCESS;
    }

    void enqueue_locked(std::unique_ptr<Req> r, uint64_t *ticket)
    {
        r->id = ++next_id_;
        if (ticket) *ticket = r->id;
        queue_.push_back(r.get());
        reqs_[r->id] = std::move(r);
    }

    void loop()
    {
        (void)hipSetDevice(device_);
        std::unique_lock<std::mutex> lk(mu_);
        while (true) {
            bool progress = retire_locked();
            progress |= admit_locked();
            if (stop_ && queue_.empty() && inflight_.empty()) break;
            if (!progress) {
                if (inflight_.empty())
                    cv_.wait_for(lk, std::chrono::milliseconds(5));
                else {
                    lk.unlock();
                    std::this_thread::sleep_for(std::chrono::microseconds(20));
                    lk.lock();
                }
            }
        }
    }

    bool retire_locked()
    {
        bool any = false;
        // launches on one stream complete in order: after the first unfinished one of a stream the
        // later ones of that stream are not queried
        std::vector<char> blocked((size_t)nstreams_, 0);
        for (size_t i = 0; i < inflight_.size();) {
            Launch &l = *inflight_[i];
            if (blocked[(size_t)l.stream]) {
                i++;
                continue;
            }
            const hipError_t q = hipEventQuery(l.ev);
            if (q == hipErrorNotReady) {
                blocked[(size_t)l.stream] = 1;
                i++;
                continue;
            }
            const double t = now_s();
            last_done_ = t;
            for (Req *r : l.reqs) {
                r->t_done = t;
                r->done = true;
                r->status = (q == hipSuccess) ? AsdSip::ACL_SUCCESS : AsdSip::ACL_ERROR_INTERNAL_ERROR;
                lat_ms_.push_back((r->t_done - r->t_arr) * 1e3);
                bytes_done_ += (double)r->batch * (r->type == kTypeC2C ? 16.0 * r->n : 4.0 * r->n + 8.0 * (r->n / 2 + 1));
                flops_done_ += (double)r->batch * (r->type == kTypeC2C ? 5.0 : 2.5) * r->n * std::log2((double)r->n);
                if (r->own_out && r->out_off != SIZE_MAX) arena_->release(r->out_off, r->out_bytes);
                if (r->status == AsdSip::ACL_SUCCESS)
                    completed_ok_++;
                else
                    completed_err_++;
                if (r->detached) reqs_.erase(r->id);
            }
            if (l.ws_off != SIZE_MAX) arena_->release(l.ws_off, l.ws_bytes);
            if (l.tab_off != SIZE_MAX) arena_->release(l.tab_off, l.tab_bytes);
            pin_release(l.pin_off, l.pin_bytes);
            inflight_per_stream_[(size_t)l.stream]--;
            ev_pool_.push_back(l.ev);  // reused by a later launch
            inflight_.erase(inflight_.begin() + (std::ptrdiff_t)i);
            any = true;
        }
        if (any) done_cv_.notify_all();
        return any;
    }

    // pinned staging ring for row tables (released in launch order)
    bool pin_alloc(size_t bytes, size_t *off)
    {
        bytes = (bytes + 255) / 256 * 256;
        if (bytes > pin_bytes_) return false;
        // ring: allocate at head if the span up to the oldest live allocation is free
        if (pin_live_.empty()) {
            pin_head_ = 0;
        }
        size_t head = pin_head_;
        if (head + bytes > pin_bytes_) head = 0;
        for (auto &kv : pin_live_) {
            const size_t a = kv.first, b = kv.first + kv.second;
            if (head < b && a < head + bytes) return false;
        }
        pin_live_[head] = bytes;
        pin_head_ = head + bytes;
        *off = head;
        return true;
    }
    void pin_release(size_t off, size_t bytes)
    {
        if (bytes) pin_live_.erase(off);
    }

    bool admit_locked()
    {
        bool any = false;
        while (!queue_.empty()) {
            // a free stream slot
            int sidx = -1;
            int best = 1 << 30;
            for (int s = 0; s < nstreams_; s++)
                if (inflight_per_stream_[(size_t)s] < 2 && inflight_per_stream_[(size_t)s] < best) {
                    best = inflight_per_stream_[(size_t)s];
                    sidx = s;
                }
            if (sidx < 0) break;
            Req *head = queue_.front();
            // coalesce queued requests of the same shape (FIFO order, bounded scan), or, for short
            // C2C lengths, of any short C2C length (one persistent multi-length launch)
            std::vector<Req *> group;
            int64_t rows = 0;
            const bool mixed = mixed_ok(head);
            const bool gathered = gathered_supported(head->type, head->n);
            size_t scanned = 0;
            for (auto it = queue_.begin(); it != queue_.end() && scanned < 256; ++it, ++scanned) {
                Req *r = *it;
                if (mixed ? !mixed_ok(r) : (r->type != head->type || r->n != head->n)) continue;
                if (!group.empty() && ((!gathered && !mixed) || rows + r->batch > max_rows_)) break;
                group.push_back(r);
                rows += r->batch;
                if (!gathered && !mixed) break;
            }
            bool ok = mixed ? launch_mixed_locked(group, sidx) : launch_locked(group, rows, sidx);
            if (!ok && group.size() > 1) {
                // the coalesced launch does not fit now: try the head request alone
                group.resize(1);
                rows = head->batch;
                ok = mixed ? launch_mixed_locked(group, sidx) : launch_locked(group, rows, sidx);
            }
            if (!ok) {
                if (!inflight_.empty()) break;  // out of memory: wait for completions to free some
                // Nothing runs, so the arena is empty and the head request can never fit (submit
                // admits only requests whose lone launch fits the budget; this catches the rest):
                // it fails instead of blocking the queue.
                head->done = true;
                head->status = AsdSip::ACL_ERROR_INTERNAL_ERROR;
                head->t_done = now_s();
                queue_.pop_front();
                completed_err_++;
                if (head->detached) reqs_.erase(head->id);
                done_cv_.notify_all();
                any = true;
                continue;
            }
            for (Req *r : group) {
                queue_.erase(std::find(queue_.begin(), queue_.end(), r));
                // a launch that failed has already completed its requests (with an error status)
                if (r->done && r->detached) reqs_.erase(r->id);
            }
            any = true;
        }
        return any;
    }

    bool launch_locked(const std::vector<Req *> &group, int64_t rows, int sidx)
    {
        Req *h = group.front();
        const int type = h->type;
        const int64_t n = h->n;
        auto l = std::make_unique<Launch>();
        l->stream = sidx;
        // outputs owned by the server
        std::vector<size_t> taken;
        auto rollback = [&]() {
            for (Req *r : group)
                if (r->own_out && r->out_off != SIZE_MAX) {
                    arena_->release(r->out_off, r->out_bytes);
                    r->out_off = SIZE_MAX;
                    r->out = nullptr;
                }
            if (l->ws_off != SIZE_MAX) arena_->release(l->ws_off, l->ws_bytes);
            if (l->tab_off != SIZE_MAX) arena_->release(l->tab_off, l->tab_bytes);
            if (l->pin_bytes) pin_release(l->pin_off, l->pin_bytes);
        };
        for (Req *r : group) {
            if (!r->own_out) continue;
            r->out_off = arena_->alloc(r->out_bytes);
            if (r->out_off == SIZE_MAX) {
                rollback();
                return false;
            }
            r->out = static_cast<char *>(arena_base_) + r->out_off;
        }
        auto plan = std::make_unique<Plan>();
        plan->direction = kDirForward;
        // a request's plan must not depend on how many rows it shares a launch with (results are
        // bit-identical alone and coalesced), so no row-count dependent single-pass leaves
        plan->big_leaf = false;
        // one request: its rows are contiguous, so the plain plan runs it without row tables (no
        // table staging, no host-to-device copy); the same kernels as the gathered plan
        const bool gathered = gathered_supported(type, n) && group.size() > 1;
        try {
            if (gathered) {
                // row tables relative to the lowest input / output address of the group
                uintptr_t in_lo = UINTPTR_MAX, out_lo = UINTPTR_MAX;
                for (Req *r : group) {
                    in_lo = std::min(in_lo, reinterpret_cast<uintptr_t>(r->in));
                    out_lo = std::min(out_lo, reinterpret_cast<uintptr_t>(r->out));
                }
                const size_t in_esz = type == kTypeC2C ? 8 : 4;
                const int64_t in_len = n, out_len = type == kTypeC2C ? n : n / 2 + 1;
                std::vector<int64_t> ir, orow;
                ir.reserve((size_t)rows);
                orow.reserve((size_t)rows);
                for (Req *r : group) {
                    const int64_t ib = (int64_t)((reinterpret_cast<uintptr_t>(r->in) - in_lo) / in_esz);
                    const int64_t ob = (int64_t)((reinterpret_cast<uintptr_t>(r->out) - out_lo) / 8);
                    for (int64_t k = 0; k < r->batch; k++) {
                        ir.push_back(ib + k * in_len);
                        orow.push_back(ob + k * out_len);
                    }
                }
                plan_build_gathered(*plan, type, n, ir, orow, reinterpret_cast<void *>(in_lo),
                                    reinterpret_cast<void *>(out_lo));
                // stage the tables: pinned ring -> arena (async on the launch stream)
                const size_t tb = plan->row_tables[0].size() * 8;
                l->tab_bytes = 2 * ((tb + 255) / 256 * 256);
                l->tab_off = arena_->alloc(l->tab_bytes);
                if (l->tab_off == SIZE_MAX || !pin_alloc(l->tab_bytes, &l->pin_off)) {
                    if (l->tab_off != SIZE_MAX) {
                        arena_->release(l->tab_off, l->tab_bytes);
                        l->tab_off = SIZE_MAX;
                    }
                    rollback();
                    return false;
                }
                l->pin_bytes = l->tab_bytes;
                char *pin = static_cast<char *>(pin_) + l->pin_off;
                std::memcpy(pin, plan->row_tables[0].data(), tb);
                std::memcpy(pin + l->tab_bytes / 2, plan->row_tables[1].data(), tb);
                char *dtab = static_cast<char *>(arena_base_) + l->tab_off;
                SIP_HIP_CHECK(hipMemcpyAsync(dtab, pin, l->tab_bytes, hipMemcpyHostToDevice, streams_[(size_t)sidx]));
                plan_materialize_external(*plan, {reinterpret_cast<const int64_t *>(dtab),
                                                  reinterpret_cast<const int64_t *>(dtab + l->tab_bytes / 2)});
            } else {
                plan->fftType = type;
                plan->sizes = {n};
                plan->batch = h->batch;
                plan_build(*plan);
                plan_materialize(*plan);
            }
            keep_resident(*plan);
            if (plan->ws_bytes) {
                l->ws_bytes = plan->ws_bytes;
                l->ws_off = arena_->alloc(l->ws_bytes);
                if (l->ws_off == SIZE_MAX) {
                    rollback();
                    return false;
                }
                plan->workspace = static_cast<char *>(arena_base_) + l->ws_off;
            }
            plan->stream = streams_[(size_t)sidx];
            const double t = now_s();
            for (Req *r : group) r->t_adm = t;
            if (gathered)
                plan_execute(*plan, nullptr, nullptr);
            else
                plan_execute(*plan, h->in, h->out);
            if (!ev_pool_.empty()) {
                l->ev = ev_pool_.back();
                ev_pool_.pop_back();
            } else {
                SIP_HIP_CHECK(hipEventCreateWithFlags(&l->ev, hipEventDisableTiming));
            }
            const hipError_t er = hipEventRecord(l->ev, streams_[(size_t)sidx]);
            if (er != hipSuccess) {
                ev_pool_.push_back(l->ev);
                l->ev = nullptr;
                SIP_HIP_CHECK(er);
            }
        } catch (...) {
            // The requests fail; the caller removes them from the queue. Work that was already
            // enqueued (a table copy, the first stages) finishes before its memory is reused.
            (void)hipStreamSynchronize(streams_[(size_t)sidx]);
            rollback();
            for (Req *r : group) {
                r->done = true;
                r->status = AsdSip::ACL_ERROR_INTERNAL_ERROR;
                r->t_done = now_s();
                completed_err_++;
            }
            done_cv_.notify_all();
            return true;
        }
        l->reqs = group;
        l->plan = std::move(plan);
        inflight_per_stream_[(size_t)sidx]++;
        launches_++;
        coalesced_ += group.size();
        inflight_.push_back(std::move(l));
        arena_peak_ = std::max(arena_peak_, arena_->used());
        return true;
    }

    // Short C2C requests (2,3,5-smooth n up to 4096) run in persistent multi-length launches.
    bool mixed_ok(const Req *r)
    {
        if (!mixed_enabled_ || r->type != kTypeC2C || r->n > kPersistMaxL) return false;
        auto it = mixed_shape_.find(r->n);
        if (it == mixed_shape_.end()) it = mixed_shape_.emplace(r->n, !persist_default_radices(r->n).empty()).first;
        return it->second;
    }

    // One persistent launch over the group's requests (their lengths may differ): outputs the server
    // owns from the arena, the launch image (descriptors, tile list, row tables) staged through the
    // pinned ring and copied to the arena on the launch stream. False when memory is short now.
    bool launch_mixed_locked(const std::vector<Req *> &group, int sidx)
    {
        auto l = std::make_unique<Launch>();
        l->stream = sidx;
        auto rollback = [&]() {
            for (Req *r : group)
                if (r->own_out && r->out_off != SIZE_MAX) {
                    arena_->release(r->out_off, r->out_bytes);
                    r->out_off = SIZE_MAX;
                    r->out = nullptr;
                }
            if (l->tab_off != SIZE_MAX) arena_->release(l->tab_off, l->tab_bytes);
            if (l->pin_bytes) pin_release(l->pin_off, l->pin_bytes);
        };
        for (Req *r : group) {
            if (!r->own_out) continue;
            r->out_off = arena_->alloc(r->out_bytes);
            if (r->out_off == SIZE_MAX) {
                rollback();
                return false;
            }
            r->out = static_cast<char *>(arena_base_) + r->out_off;
        }
        try {
            uintptr_t in_lo = UINTPTR_MAX, out_lo = UINTPTR_MAX;
            for (Req *r : group) {
                in_lo = std::min(in_lo, reinterpret_cast<uintptr_t>(r->in));
                out_lo = std::min(out_lo, reinterpret_cast<uintptr_t>(r->out));
            }
            std::map<int64_t, PersistSpec> by;
            for (Req *r : group) {
                PersistSpec &sp = by[r->n];
                sp.n = r->n;
                const int64_t ib = (int64_t)((reinterpret_cast<uintptr_t>(r->in) - in_lo) / 8);
                const int64_t ob = (int64_t)((reinterpret_cast<uintptr_t>(r->out) - out_lo) / 8);
                for (int64_t k = 0; k < r->batch; k++) {
                    sp.in_rows.push_back(ib + k * r->n);
                    sp.out_rows.push_back(ob + k * r->n);
                }
            }
            std::vector<PersistSpec> specs;
            specs.reserve(by.size());
            for (auto it = by.rbegin(); it != by.rend(); ++it) specs.push_back(std::move(it->second));  // longest first
            PersistImage img;
            persist_build_image(specs, reinterpret_cast<const void *>(in_lo), reinterpret_cast<void *>(out_lo), -1, img);
            l->tab_bytes = granule(img.bytes.size());
            l->tab_off = arena_->alloc(l->tab_bytes);
            if (l->tab_off == SIZE_MAX || !pin_alloc(l->tab_bytes, &l->pin_off)) {
                if (l->tab_off != SIZE_MAX) {
                    arena_->release(l->tab_off, l->tab_bytes);
                    l->tab_off = SIZE_MAX;
                }
                rollback();
                return false;
            }
            l->pin_bytes = l->tab_bytes;
            char *dimg = static_cast<char *>(arena_base_) + l->tab_off;
            persist_place(img, dimg);
            char *pin = static_cast<char *>(pin_) + l->pin_off;
            std::memcpy(pin, img.bytes.data(), img.bytes.size());
            hipStream_t st = streams_[(size_t)sidx];
            SIP_HIP_CHECK(hipMemcpyAsync(dimg, pin, img.bytes.size(), hipMemcpyHostToDevice, st));
            const double t = now_s();
            for (Req *r : group) r->t_adm = t;
            persist_launch(img, st);
            l->hold = std::move(img.tables);
            if (!ev_pool_.empty()) {
                l->ev = ev_pool_.back();
                ev_pool_.pop_back();
            } else {
                SIP_HIP_CHECK(hipEventCreateWithFlags(&l->ev, hipEventDisableTiming));
            }
            const hipError_t er = hipEventRecord(l->ev, st);
            if (er != hipSuccess) {
                ev_pool_.push_back(l->ev);
                l->ev = nullptr;
                SIP_HIP_CHECK(er);
            }
        } catch (...) {
            (void)hipStreamSynchronize(streams_[(size_t)sidx]);
            rollback();
            for (Req *r : group) {
                r->done = true;
                r->status = AsdSip::ACL_ERROR_INTERNAL_ERROR;
                r->t_done = now_s();
                completed_err_++;
            }
            done_cv_.notify_all();
            return true;
        }
        l->reqs = group;
        inflight_per_stream_[(size_t)sidx]++;
        launches_++;
        mixed_launches_++;
        coalesced_ += group.size();
        inflight_.push_back(std::move(l));
        arena_peak_ = std::max(arena_peak_, arena_->used());
        return true;
    }

    // Every coefficient table a launch has used stays on the device: a later launch reuses it
    // without a rebuild or an upload, and no table is freed (a device-synchronising hipFree) while
    // launches run. Tables are held one by one, not per shape, so that every table any launch of a
    // shape used stays (a single request and a coalesced group build separate plans). The resident
    // set is capped (resident_cap_): beyond it the least recently used tables are dropped (a launch
    // in flight keeps its own reference until it retires).
    void keep_resident(const Plan &plan)
    {
        const uint64_t now = ++resident_clock_;
        for (const Stage &st : plan.stages)
            for (const TableRef *t : {&st.tw, &st.rtw, &st.tw4lo, &st.tw4hi, &st.mfw, &st.tab0, &st.tw2}) {
                if (!t->dev) continue;
                auto ins = resident_.emplace(t->dev.get(), Resident{t->dev, now});
                if (ins.second)
                    resident_bytes_ += t->dev->bytes;
                else
                    ins.first->second.last = now;
            }
        while (resident_bytes_ > resident_cap_) {
            auto victim = resident_.end();
            for (auto it = resident_.begin(); it != resident_.end(); ++it)
                if (it->second.last != now && (victim == resident_.end() || it->second.last < victim->second.last))
                    victim = it;
            if (victim == resident_.end()) break;  // only the current launch's tables are left
            resident_bytes_ -= victim->second.table->bytes;
            resident_.erase(victim);
            resident_evictions_++;
        }
    }

    size_t budget_;
    int nstreams_;
    // Multi-length persistent launches only with SIP_FFT_SERVE_MIXED=1. Measured on MI355X (10 s
    // Poisson windows of the serving configuration): they cut launches at 50,000 req/s from 481k
    // to 341k but raised p99 (10k req/s 0.24 -> 0.88 ms, 50k 1.27 -> 2.66 ms), so single-length
    // launches stay the default.
    bool mixed_enabled_ = [] {
        const char *e = std::getenv("SIP_FFT_SERVE_MIXED");
        return e && e[0] == '1';
    }();
    std::map<int64_t, bool> mixed_shape_;  // n -> runs in persistent launches
    uint64_t mixed_launches_ = 0;
    struct Resident {
        TablePtr table;
        uint64_t last = 0;  // resident_clock_ of the latest launch that used it
    };
    std::unordered_map<const void *, Resident> resident_;
    size_t resident_bytes_ = 0, resident_cap_ = 0;
    uint64_t resident_clock_ = 0, resident_evictions_ = 0;
    int64_t max_rows_;
    int device_ = 0;
    void *arena_base_ = nullptr;
    std::unique_ptr<Arena> arena_;
    void *pin_ = nullptr;
    size_t pin_bytes_ = 0, pin_head_ = 0;
    std::map<size_t, size_t> pin_live_;
    std::vector<hipStream_t> streams_;
    std::vector<int> inflight_per_stream_;
    std::vector<std::unique_ptr<Launch>> inflight_;
    std::deque<Req *> queue_;
    std::unordered_map<uint64_t, std::unique_ptr<Req>> reqs_;
    uint64_t next_id_ = 0;
    std::vector<double> lat_ms_;
    uint64_t completed_ok_ = 0, completed_err_ = 0;
    uint64_t launches_ = 0, coalesced_ = 0;
    double bytes_done_ = 0, flops_done_ = 0;
    size_t arena_peak_ = 0;
    double last_done_ = 0;
    std::mutex mu_;
    std::condition_variable cv_, done_cv_;
    std::vector<hipEvent_t> ev_pool_;  // completed launches' events, recorded again by later launches
    bool stop_ = false;
    std::thread worker_;
};

std::mutex g_srv_mu;
std::unordered_map<void *, std::unique_ptr<Server>> g_servers;

Server *lookup_server(void *h)
{
    std::lock_guard<std::mutex> lk(g_srv_mu);
    auto it = g_servers.find(h);
    return it == g_servers.end() ? nullptr : it->second.get();
}

template <class F>
int32_t guarded(F &&f)
{
    try {
        return f();
    } catch (const SipError &e) {
        SIP_LOG(ERROR, e.what());
        return e.status;
    } catch (const std::exception &e) {
        SIP_LOG(ERROR, e.what());
        return AsdSip::ACL_ERROR_INTERNAL_ERROR;
    } catch (...) {
        return AsdSip::ACL_ERROR_INTERNAL_ERROR;
    }
}

std::vector<int64_t> smooth_between(int64_t lo, int64_t hi)
{
    std::vector<int64_t> v;
    for (int64_t a = 1; a <= hi; a *= 2)
        for (int64_t b = a; b <= hi; b *= 3)
            for (int64_t c = b; c <= hi; c *= 5)
                if (c >= lo) v.push_back(c);
    std::sort(v.begin(), v.end());
    return v;
}

}  // namespace
}  // namespace sip

namespace E = AsdSip;

extern "C" {

int32_t sipFftServeCreate(void **handle, size_t budget_bytes, int32_t nstreams, int64_t max_coalesce_rows)
{
    return sip::guarded([&]() -> int32_t {
        if (!handle || budget_bytes < (1u << 20) || nstreams < 1 || nstreams > 16 || max_coalesce_rows < 1)
            return E::ACL_ERROR_INVALID_PARAM;
        auto s = std::make_unique<sip::Server>(budget_bytes, nstreams, max_coalesce_rows);
        void *h = s.get();
        std::lock_guard<std::mutex> lk(sip::g_srv_mu);
        sip::g_servers.emplace(h, std::move(s));
        *handle = h;
        return E::ACL_SUCCESS;
    });
}

int32_t sipFftServeSubmit(void *handle, int32_t fftType, int64_t n, int64_t batch, const void *input, void *output,
                          uint64_t *ticket)
{
    return sip::guarded([&]() -> int32_t {
        sip::Server *s = sip::lookup_server(handle);
        if (!s) return E::ACL_ERROR_INVALID_PARAM;
        return s->submit(fftType, n, batch, input, output, ticket);
    });
}

int32_t sipFftServeSubmitMany(void *handle, int64_t count, const int32_t *fftTypes, const int64_t *n,
                              const int64_t *batch, const void *const *inputs, void *const *outputs,
                              uint64_t *tickets, int64_t *refused)
{
    return sip::guarded([&]() -> int32_t {
        sip::Server *s = sip::lookup_server(handle);
        if (!s) return E::ACL_ERROR_INVALID_PARAM;
        return s->submit_many(count, fftTypes, n, batch, inputs, outputs, tickets, refused);
    });
}

int32_t sipFftServeWait(void *handle, uint64_t ticket, double *latency_ms)
{
    return sip::guarded([&]() -> int32_t {
        sip::Server *s = sip::lookup_server(handle);
        if (!s) return E::ACL_ERROR_INVALID_PARAM;
        return s->wait(ticket, latency_ms);
    });
}

int32_t sipFftServeDrain(void *handle)
{
    return sip::guarded([&]() -> int32_t {
        sip::Server *s = sip::lookup_server(handle);
        if (!s) return E::ACL_ERROR_INVALID_PARAM;
        s->drain();
        return E::ACL_SUCCESS;
    });
}

int32_t sipFftServeStats(void *handle, int32_t reset, char *buf, size_t cap, size_t *needed)
{
    return sip::guarded([&]() -> int32_t {
        sip::Server *s = sip::lookup_server(handle);
        if (!s) return E::ACL_ERROR_INVALID_PARAM;
        const std::string j = s->stats(reset != 0, buf ? cap : 0);
        if (needed) *needed = j.size() + 1;
        if (!buf || cap < j.size() + 1) return E::ACL_ERROR_INVALID_PARAM;
        std::memcpy(buf, j.c_str(), j.size() + 1);
        return E::ACL_SUCCESS;
    });
}

// Open-loop load generator: Poisson arrivals at `rate` req/s for `seconds`; each request is C2C
// with probability `c2c_frac` (else R2C), length log-uniform over the 2,3,5-smooth numbers in
// [n_lo, n_hi], batch uniform in [1, b_max]; its input is a random slice of the device pool
// (uniform[-1,1) data prepared by the caller). Outputs come from the server's arena. Waits for the
// queue to drain (at most `drain_seconds`), then writes the server statistics plus offered/achieved
// rates into buf.
int32_t sipFftServeBenchmark(void *handle, double rate, double seconds, uint64_t seed, const void *pool,
                             size_t pool_bytes, double c2c_frac, int64_t n_lo, int64_t n_hi, int64_t b_max,
                             double drain_seconds, char *buf, size_t cap, size_t *needed)
{
    return sip::guarded([&]() -> int32_t {
        sip::Server *s = sip::lookup_server(handle);
        if (!s || rate <= 0 || seconds <= 0 || !pool || n_lo < 1 || n_hi < n_lo || b_max < 1)
            return E::ACL_ERROR_INVALID_PARAM;
        const std::vector<int64_t> sm = sip::smooth_between(n_lo, n_hi);
        if (sm.empty()) return E::ACL_ERROR_INVALID_PARAM;
        // every request's input (one row at least) must fit in the pool
        if (pool_bytes < (size_t)sm.back() * 8) return E::ACL_ERROR_INVALID_PARAM;
        std::mt19937_64 rng(seed);
        std::exponential_distribution<double> gap(rate);
        std::uniform_real_distribution<double> u01(0.0, 1.0);
        std::uniform_int_distribution<int64_t> ub(1, b_max);
        // pre-generate the request stream
        struct Gen {
            double t;
            int type;
            int64_t n, b;
            size_t off;
        };
        std::vector<Gen> gen;
        double t = 0;
        const double lo = std::log((double)n_lo), hi = std::log((double)n_hi);
        while (true) {
            t += gap(rng);
            if (t >= seconds) break;
            Gen g;
            g.t = t;
            g.type = u01(rng) < c2c_frac ? sip::kTypeC2C : sip::kTypeR2C;
            const double x = std::exp(lo + (hi - lo) * u01(rng));
            g.n = *std::lower_bound(sm.begin(), sm.end(), (int64_t)std::ceil(x));
            g.b = ub(rng);
            const size_t esz = g.type == sip::kTypeC2C ? 8 : 4;
            if ((size_t)g.b * (size_t)g.n * esz > pool_bytes)
                g.b = std::max<int64_t>(1, (int64_t)(pool_bytes / ((size_t)g.n * esz)));
            const size_t need2 = (size_t)g.b * (size_t)g.n * esz;  // <= pool_bytes (checked above)
            const size_t span = pool_bytes - need2;
            g.off = (size_t)(u01(rng) * (double)(span / 8)) * 8;
            gen.push_back(g);
        }
        (void)s->stats(true);
        s->drain();
        const auto c0 = s->completed();
        const double t0 = sip::now_s();
        // requests are detached: released as they complete (latencies still enter the stats)
        for (const Gen &g : gen) {
            const double due = t0 + g.t;
            double now = sip::now_s();
            if (due > now) std::this_thread::sleep_for(std::chrono::duration<double>(due - now));
            const int32_t st =
                s->submit(g.type, g.n, g.b, static_cast<const char *>(pool) + g.off, nullptr, nullptr, true);
            if (st != E::ACL_SUCCESS) return st;
        }
        const double t_sub = sip::now_s();
        // Serve until the queue drains or the drain window closes; what is still queued then is
        // dropped (an overloaded stream would otherwise grow without bound), in-flight work finishes.
        const double limit = t0 + seconds + drain_seconds;
        while (sip::now_s() < limit && !s->idle()) std::this_thread::sleep_for(std::chrono::milliseconds(1));
        const size_t dropped = s->cancel_queued();
        s->drain();
        const auto c1 = s->completed();
        const size_t ok = (size_t)(c1.first - c0.first);
        const int32_t worst = c1.second > c0.second ? E::ACL_ERROR_INTERNAL_ERROR : E::ACL_SUCCESS;
        const double t_last = s->last_done();
        std::string st = s->stats(false);
        std::ostringstream os;
        os << "{\"offered_rate\":" << rate << ",\"seconds\":" << seconds << ",\"submitted\":" << gen.size()
           << ",\"completed\":" << ok << ",\"dropped_unadmitted\":" << dropped << ",\"busy_s\":" << (t_last - t0)
           << ",\"achieved_req_per_s\":" << (double)ok / std::max(1e-9, t_last - t0)
           << ",\"submit_window_s\":" << (t_sub - t0) << ",\"status\":" << worst << ",\"server\":" << st << "}";
        const std::string j = os.str();
        if (needed) *needed = j.size() + 1;
        if (!buf || cap < j.size() + 1) return E::ACL_ERROR_INVALID_PARAM;
        std::memcpy(buf, j.c_str(), j.size() + 1);
        return E::ACL_SUCCESS;
    });
}

int32_t sipFftServeDestroy(void *handle)
{
    std::unique_ptr<sip::Server> victim;
    {
        std::lock_guard<std::mutex> lk(sip::g_srv_mu);
        auto it = sip::g_servers.find(handle);
        if (it == sip::g_servers.end()) return E::ACL_ERROR_INVALID_PARAM;
        victim = std::move(it->second);
        sip::g_servers.erase(it);
    }
    victim.reset();
    return E::ACL_SUCCESS;
}

}  // extern "C"
