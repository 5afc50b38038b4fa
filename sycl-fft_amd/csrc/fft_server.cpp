// Serving loop of the MI355X provider: continuous batching of transform requests over
// resident plans (C ABI in include/syclfft_hip_server.h).
//
//   submit()      validates a request, queues it, returns a ticket
//   dispatcher    takes the oldest waiting request, merges the other waiting requests of the
//                 same shape into its launch (up to the per-launch transform limit), admits
//                 the group if the HBM footprint (request buffers in flight + resident
//                 plans) stays within the budget and a launch slot is free, picks or creates
//                 a plan whose batch count covers the group, and enqueues one pointer-array
//                 execution on the next stream
//   completer     polls the launches' events, completes their requests (latency records),
//                 returns their footprint and wakes the dispatcher and waiters
//
// Plans stay resident between requests (twiddle tables, chirp spectra, scratch) and are
// evicted least recently used when a new plan would not fit the budget.  A plan is created
// only when no resident plan of the shape can hold the oldest waiting request (creation
// costs far more than a launch); requests of different shapes admitted while earlier
// launches run go to the least loaded stream, so their work overlaps on the device.  No reference counterpart: the reference's only
// concurrency model is one plan per concurrent stream of work (README.md:45-47).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <list>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "syclfft_hip_server.h"

// fft_plan.cpp: pointer-array execution for buffers validated as whole allocations
extern "C" int sfh_plan_execute_many_checked_buffers(
    void *plan, uint64_t count, const void *const *ins, void *const *outs, void *stream, void *done, char *err,
    size_t cap);
// fft_plan.cpp: work-queue class of `count` transforms of a plan (-1: none) and a queue
// execution of checked buffers (entry k: counts[k] transforms of plans[k])
extern "C" int sfh_plan_queue_class(const void *plan, uint64_t count);
// fft_plan.cpp: device memory a plan holds
extern "C" uint64_t sfh_plan_device_bytes(const void *plan);
extern "C" int sfh_execute_queue_checked_buffers(
    void *const *plans, const uint64_t *counts, const void *const *ins, void *const *outs, size_t n, void *stream,
    void *done, char *err, size_t cap);

namespace sfh_srv
{

static int64_t now_ns()
{
    return std::chrono::duration_cast<std::chrono::nanoseconds>(std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

struct srv_error : std::runtime_error
{
    int status;
    srv_error(int s, const std::string &m)
      : std::runtime_error(m), status(s)
    {}
};

static void set_err(char *err, size_t cap, const std::string &msg)
{
    if (err && cap) {
        const size_t n = std::min(cap - 1, msg.size());
        std::memcpy(err, msg.data(), n);
        err[n] = '\0';
    }
}

struct request
{
    syclfft_hip_request_v1 r{};
    uint64_t ticket{};
    std::string key;
    uint64_t bytes{}; // HBM footprint of its buffers
    int64_t t_submit{}, t_launch{}, t_done{};
    int status{SYCLFFT_HIP_STATUS_OK};
    std::string err;
    uint32_t group{};
    bool done{false};
};

struct plan_entry
{
    std::string key;
    void *plan{};
    uint64_t capacity{}; // transforms one launch may hold (the plan's batch count)
    uint64_t bytes{};    // device memory charged to the budget
    int inflight{};      // launches in flight on this plan
    int64_t last_use{};
};

struct launch_rec
{
    std::vector<std::shared_ptr<request>> reqs;
    std::vector<plan_entry *> pes; // plans the launch runs (one per request in queue launches)
    hipEvent_t done{};
    uint64_t bytes{};
    size_t stream{};
};

static uint64_t elems_of(const syclfft_hip_request_v1 &r)
{
    uint64_t n = 1;
    for (uint32_t i = 0; i < r.rank; ++i) {
        n *= r.lengths[i];
    }
    return n;
}

class server
{
public:
    server(int device, uint64_t budget, uint32_t nstreams, uint64_t max_batch)
      : budget_(budget), max_batch_(max_batch ? max_batch : 512)
    {
        if (device < 0) {
            if (hipGetDevice(&device) != hipSuccess) {
                throw srv_error(SYCLFFT_HIP_STATUS_PROVIDER_UNAVAILABLE, "no current HIP device");
            }
        }
        char why[512];
        if (!syclfft_hip_supports(device, why, sizeof(why))) {
            throw srv_error(SYCLFFT_HIP_STATUS_PROVIDER_UNAVAILABLE, why);
        }
        if (nstreams < 1 || nstreams > 16) {
            throw srv_error(SYCLFFT_HIP_STATUS_INVALID_ARGUMENT, "streams must be 1..16");
        }
        if (budget == 0) {
            throw srv_error(SYCLFFT_HIP_STATUS_INVALID_ARGUMENT, "memory budget must be positive");
        }
        device_ = device;
        if (const char *q = std::getenv("SYCLFFT_HIP_SERVE_QUEUE")) {
            queue_launches_ = std::atoi(q) != 0;
        }
        if (const char *q = std::getenv("SYCLFFT_HIP_SERVE_SJF")) {
            sjf_ = std::atoi(q) != 0;
        }
        if (const char *q = std::getenv("SYCLFFT_HIP_SERVE_SJF_AGE_US")) {
            sjf_age_ns_ = 1000 * static_cast<int64_t>(std::strtoll(q, nullptr, 10));
        }
        if (hipSetDevice(device_) != hipSuccess) {
            throw srv_error(SYCLFFT_HIP_STATUS_PROVIDER_UNAVAILABLE, "hipSetDevice failed");
        }
        for (uint32_t i = 0; i < nstreams; ++i) {
            hipStream_t s{};
            if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) {
                release_streams();
                throw srv_error(SYCLFFT_HIP_STATUS_EXECUTION_FAILED, "hipStreamCreate failed");
            }
            streams_.push_back(s);
        }
        // launches in flight before waiting requests stay on the host, where they can merge
        // (SYCLFFT_HIP_SERVE_INFLIGHT overrides 4 per stream)
        max_inflight_ = 4 * nstreams;
        if (const char *m = std::getenv("SYCLFFT_HIP_SERVE_INFLIGHT")) {
            max_inflight_ = std::max<size_t>(1, std::strtoul(m, nullptr, 10));
        }
        stream_load_.assign(nstreams, 0);
        dispatcher_ = std::thread([this] { dispatch_loop(); });
        completer_ = std::thread([this] { complete_loop(); });
    }

    ~server()
    {
        drain();
        {
            std::lock_guard<std::mutex> lk(mu_);
            stop_ = true;
        }
        cv_dispatch_.notify_all();
        cv_complete_.notify_all();
        dispatcher_.join();
        completer_.join();
        (void)hipSetDevice(device_);
        for (auto &pe : plans_) {
            syclfft_hip_plan_destroy(pe->plan);
        }
        for (auto e : events_) {
            (void)hipEventDestroy(e);
        }
        release_streams();
    }

    uint64_t submit(const syclfft_hip_request_v1 &r)
    {
        validate(r);
        auto q = std::make_shared<request>();
        q->r = r;
        const uint64_t es = r.scalar == SYCLFFT_HIP_SCALAR_DOUBLE ? 16 : 8;
        const uint64_t data = elems_of(r) * r.batch_count * es;
        check_buffer(r.in, data, "input");
        if (!r.in_place) {
            check_buffer(r.out, data, "output");
        }
        q->bytes = r.in_place ? data : 2 * data;
        if (q->bytes > budget_) {
            throw srv_error(SYCLFFT_HIP_STATUS_INVALID_ARGUMENT, "request footprint exceeds the memory budget");
        }
        char key[256];
        std::snprintf(key, sizeof(key), "%u|%d|%u|%llux%llux%llu|%u|%u", r.scalar, r.direction, r.rank,
                      static_cast<unsigned long long>(r.lengths[0]),
                      static_cast<unsigned long long>(r.rank > 1 ? r.lengths[1] : 1),
                      static_cast<unsigned long long>(r.rank > 2 ? r.lengths[2] : 1), r.normalization, r.in_place);
        q->key = key;
        q->t_submit = now_ns();
        uint64_t t;
        {
            std::lock_guard<std::mutex> lk(mu_);
            t = next_ticket_++;
            q->ticket = t;
            pending_.push_back(q);
            table_[t] = q;
            ++outstanding_;
            ++st_.submitted;
        }
        cv_dispatch_.notify_one();
        return t;
    }

    int wait(uint64_t ticket, syclfft_hip_completion_v1 *out, std::string &err)
    {
        std::unique_lock<std::mutex> lk(mu_);
        auto it = table_.find(ticket);
        if (it == table_.end()) {
            throw srv_error(SYCLFFT_HIP_STATUS_INVALID_ARGUMENT, "unknown or already waited ticket");
        }
        std::shared_ptr<request> q = it->second;
        cv_done_.wait(lk, [&] { return q->done; });
        table_.erase(ticket);
        if (out) {
            out->ticket = q->ticket;
            out->tag = q->r.tag;
            out->status = q->status;
            out->group_size = q->group;
            const int64_t tl = q->t_launch ? q->t_launch : q->t_done;
            out->queue_seconds = (tl - q->t_submit) * 1e-9;
            out->run_seconds = (q->t_done - tl) * 1e-9;
            out->latency_seconds = (q->t_done - q->t_submit) * 1e-9;
            out->submit_at = (q->t_submit - t0_) * 1e-9;
            out->launch_at = (tl - t0_) * 1e-9;
            out->done_at = (q->t_done - t0_) * 1e-9;
        }
        err = q->err;
        return q->status;
    }

    void drain()
    {
        std::unique_lock<std::mutex> lk(mu_);
        cv_done_.wait(lk, [&] { return outstanding_ == 0; });
    }

    void stats(syclfft_hip_server_stats_v1 *s)
    {
        std::lock_guard<std::mutex> lk(mu_);
        *s = st_;
        s->plans_resident = plans_.size();
        s->budget_bytes = budget_;
    }

private:
    // completion events are recycled (creating one per launch costs more than the launch)
    hipEvent_t take_event()
    {
        {
            std::lock_guard<std::mutex> lk(ev_mu_);
            if (!events_.empty()) {
                hipEvent_t e = events_.back();
                events_.pop_back();
                return e;
            }
        }
        hipEvent_t e = nullptr;
        return hipEventCreateWithFlags(&e, hipEventDisableTiming) == hipSuccess ? e : nullptr;
    }

    void give_event(hipEvent_t e)
    {
        std::lock_guard<std::mutex> lk(ev_mu_);
        events_.push_back(e);
    }

    void release_streams()
    {
        for (auto s : streams_) {
            (void)hipStreamDestroy(s);
        }
        streams_.clear();
    }

    static void validate(const syclfft_hip_request_v1 &r)
    {
        if (r.abi_version != SYCLFFT_HIP_SERVER_ABI_VERSION) {
            throw srv_error(SYCLFFT_HIP_STATUS_PLUGIN_ERROR, "request ABI version mismatch");
        }
        if (r.scalar != SYCLFFT_HIP_SCALAR_FLOAT && r.scalar != SYCLFFT_HIP_SCALAR_DOUBLE) {
            throw srv_error(SYCLFFT_HIP_STATUS_INVALID_ARGUMENT, "scalar must be float or double");
        }
        if (r.direction != -1 && r.direction != 1) {
            throw srv_error(SYCLFFT_HIP_STATUS_INVALID_ARGUMENT, "direction must be -1 or +1");
        }
        if (r.rank < 1 || r.rank > 3) {
            throw srv_error(SYCLFFT_HIP_STATUS_INVALID_ARGUMENT, "rank must be 1..3");
        }
        for (uint32_t i = 0; i < r.rank; ++i) {
            if (r.lengths[i] == 0) {
                throw srv_error(SYCLFFT_HIP_STATUS_INVALID_ARGUMENT, "lengths must be positive");
            }
        }
        if (r.batch_count == 0) {
            throw srv_error(SYCLFFT_HIP_STATUS_INVALID_ARGUMENT, "batch count must be positive");
        }
        if (r.normalization > SYCLFFT_HIP_NORM_ORTHOGONAL) {
            throw srv_error(SYCLFFT_HIP_STATUS_INVALID_ARGUMENT, "unknown normalization");
        }
        if (!r.in || !r.out) {
            throw srv_error(SYCLFFT_HIP_STATUS_INVALID_POINTER, "request buffers must be non-null");
        }
        if ((r.in == r.out) != (r.in_place != 0)) {
            throw srv_error(SYCLFFT_HIP_STATUS_INVALID_ARGUMENT,
                            r.in_place ? "in-place request needs in == out" : "out-of-place request needs in != out");
        }
    }

    void check_buffer(const void *p, uint64_t bytes, const char *name) const
    {
        hipPointerAttribute_t attr{};
        if (hipPointerGetAttributes(&attr, p) != hipSuccess) {
            (void)hipGetLastError();
            throw srv_error(SYCLFFT_HIP_STATUS_INVALID_POINTER, std::string(name) + " is not a HIP device allocation");
        }
        if (attr.type != hipMemoryTypeDevice || attr.device != device_) {
            throw srv_error(SYCLFFT_HIP_STATUS_INVALID_POINTER, std::string(name) + " is not memory of the server's device");
        }
        hipDeviceptr_t base = nullptr;
        size_t size = 0;
        if (hipMemGetAddressRange(&base, &size, const_cast<void *>(p)) == hipSuccess && base) {
            const auto b = reinterpret_cast<uintptr_t>(base), q = reinterpret_cast<uintptr_t>(p);
            if (q < b || q - b + bytes > size) {
                throw srv_error(SYCLFFT_HIP_STATUS_INVALID_POINTER, std::string(name) + " allocation is smaller than the request");
            }
        } else {
            (void)hipGetLastError();
        }
    }

    // Drop resident plans that no launch uses, least recently used first, until `need` more
    // bytes fit the budget.  Returns whether they fit.
    bool make_room(uint64_t need)
    {
        while (footprint_ + need > budget_) {
            auto victim = plans_.end();
            for (auto it = plans_.begin(); it != plans_.end(); ++it) {
                if ((*it)->inflight == 0 && (victim == plans_.end() || (*it)->last_use < (*victim)->last_use)) {
                    victim = it;
                }
            }
            if (victim == plans_.end()) {
                return false;
            }
            syclfft_hip_plan_destroy((*victim)->plan);
            footprint_ -= (*victim)->bytes;
            auto &same = by_key_[(*victim)->key];
            same.erase(std::find(same.begin(), same.end(), victim->get()));
            plans_.erase(victim);
            ++st_.plans_evicted;
        }
        return true;
    }

    void fail(const std::vector<std::shared_ptr<request>> &g, int status, const std::string &msg)
    {
        const int64_t t = now_ns();
        for (auto &q : g) {
            q->status = status;
            q->err = msg;
            q->t_done = t;
            q->group = static_cast<uint32_t>(g.size());
            q->done = true;
            --outstanding_;
            ++st_.failed;
        }
        cv_done_.notify_all();
    }

    // Admit and launch the group of the oldest waiting request; false if it has to wait
    // for a completion (budget or slots).
    bool launch_next(std::unique_lock<std::mutex> &lk)
    {
        // The oldest request leads the next launch; with size-first dispatch
        // (SYCLFFT_HIP_SERVE_SJF=1) the smallest of the first 64 waiting requests does, until
        // the oldest has waited sjf_age_ns_ (no request waits behind smaller ones for longer)
        std::shared_ptr<request> head = pending_.front();
        if (sjf_ && now_ns() - head->t_submit < sjf_age_ns_) {
            size_t looked = 0;
            for (auto it = pending_.begin(); it != pending_.end() && looked < 64; ++it, ++looked) {
                if ((*it)->bytes < head->bytes) {
                    head = *it;
                }
            }
        }
        const uint64_t cap = std::max<uint64_t>(max_batch_, head->r.batch_count);
        std::vector<std::shared_ptr<request>> group{head};
        uint64_t units = head->r.batch_count, bytes = head->bytes;
        size_t scanned = 0;
        for (auto it = pending_.begin(); it != pending_.end() && scanned < 1024; ++it, ++scanned) {
            const auto &q = *it;
            if (q == head) {
                continue;
            }
            if (q->key == head->key && units + q->r.batch_count <= cap && bytes + q->bytes <= budget_ / 2) {
                group.push_back(q);
                units += q->r.batch_count;
                bytes += q->bytes;
            }
        }
        // Plan creation (table generation and upload, allocation) costs far more than a
        // launch, so a group larger than every resident plan of its shape is trimmed to the
        // largest one that still holds the oldest request; a new plan is made only when none
        // does.
        auto &same = by_key_[head->key];
        uint64_t largest = 0;
        for (plan_entry *e : same) {
            largest = std::max(largest, e->capacity);
        }
        if (largest >= head->r.batch_count && units > largest) {
            std::vector<std::shared_ptr<request>> kept{head};
            units = head->r.batch_count;
            bytes = head->bytes;
            for (size_t k = 1; k < group.size(); ++k) {
                if (units + group[k]->r.batch_count <= largest) {
                    units += group[k]->r.batch_count;
                    bytes += group[k]->bytes;
                    kept.push_back(group[k]);
                }
            }
            group.swap(kept);
        }
        // an idle resident plan of this shape large enough for the group, else the least busy
        // of them (its executions serialize; waiting requests of the shape merge into the next
        // launch instead), else a new one
        plan_entry *pe = nullptr, *busy = nullptr;
        for (plan_entry *e : same) {
            if (e->capacity < units) {
                continue;
            }
            if (e->inflight == 0 && (!pe || e->capacity < pe->capacity)) {
                pe = e;
            }
            if (!busy || e->inflight < busy->inflight) {
                busy = e;
            }
        }
        if (!pe && busy) {
            pe = busy;
        }
        uint64_t capacity = 0, plan_bytes = 0;
        if (!pe) {
            capacity = 8;
            while (capacity < units) {
                capacity *= 2;
            }
            capacity = std::min(capacity, cap);
            // charged up front: scratch of a multi-launch plan is about 2 transforms' worth per
            // transform, tables about 4 transforms' worth once
            const uint64_t es = head->r.scalar == SYCLFFT_HIP_SCALAR_DOUBLE ? 16 : 8;
            plan_bytes = (2 * capacity + 4) * elems_of(head->r) * es;
        }
        // The resident plan chosen for the group is pinned while room is made, so the
        // eviction cannot destroy it (the group would run on a freed plan).
        struct pin_guard
        {
            plan_entry *e;
            explicit pin_guard(plan_entry *x) : e(x)
            {
                if (e) {
                    ++e->inflight;
                }
            }
            ~pin_guard()
            {
                if (e) {
                    --e->inflight;
                }
            }
        } pin(pe);
        if (!make_room(bytes + plan_bytes)) {
            if (!inflight_.empty()) {
                return false; // wait for completions to free memory
            }
            // nothing in flight: the group can never fit as it is; shrink it to the oldest
            // request, and fail that one if even it does not fit
            if (group.size() > 1) {
                group.resize(1);
                units = head->r.batch_count;
                bytes = head->bytes;
                if (!pe) {
                    capacity = 8;
                    while (capacity < units) {
                        capacity *= 2;
                    }
                    capacity = std::min(capacity, cap);
                    const uint64_t es = head->r.scalar == SYCLFFT_HIP_SCALAR_DOUBLE ? 16 : 8;
                    plan_bytes = (2 * capacity + 4) * elems_of(head->r) * es;
                }
            }
            if (group.size() > 1 || !make_room(bytes + plan_bytes)) {
                pending_.erase(std::find(pending_.begin(), pending_.end(), head));
                fail(group, SYCLFFT_HIP_STATUS_INVALID_ARGUMENT, "request and its plan exceed the memory budget");
                return true;
            }
        }
        // Work queue: waiting requests of other shapes whose resident plans share the head
        // plan's work-queue class (one plain tile launch of the same kernel variant) join the
        // launch, which then runs every request's tiles from one persistent launch.
        std::vector<plan_entry *> entry_plans; // per request, when the launch is a queue launch
        if (pe && queue_launches_) {
            int cls = -1;
            bool all = true;
            for (auto &q : group) {
                const int c = sfh_plan_queue_class(pe->plan, q->r.batch_count);
                all = all && c >= 0 && (cls < 0 || c == cls);
                cls = c;
            }
            std::vector<std::shared_ptr<request>> extra;
            std::vector<plan_entry *> extra_pe;
            size_t looked = 0;
            for (auto it = pending_.begin(); all && it != pending_.end() && looked < 1024 && extra.size() < 255;
                 ++it, ++looked) {
                const auto &q = *it;
                if (q->key == head->key || footprint_ + bytes + q->bytes > budget_ || bytes + q->bytes > budget_ / 2) {
                    continue; // same shape: in the group already, or beyond its plan's capacity
                }
                const auto f = by_key_.find(q->key);
                if (f == by_key_.end()) {
                    continue;
                }
                plan_entry *best = nullptr;
                for (plan_entry *e : f->second) {
                    if (e->capacity >= q->r.batch_count && (!best || e->inflight < best->inflight)) {
                        best = e;
                    }
                }
                if (best && sfh_plan_queue_class(best->plan, q->r.batch_count) == cls) {
                    extra.push_back(q);
                    extra_pe.push_back(best);
                    bytes += q->bytes;
                }
            }
            if (!extra.empty()) {
                entry_plans.assign(group.size(), pe);
                for (size_t k = 0; k < extra.size(); ++k) {
                    group.push_back(extra[k]);
                    entry_plans.push_back(extra_pe[k]);
                }
            }
        }
        // reserve: take the group out of the queue, charge its footprint
        for (auto &q : group) {
            pending_.erase(std::find(pending_.begin(), pending_.end(), q));
        }
        footprint_ += bytes + plan_bytes;
        st_.peak_footprint_bytes = std::max(st_.peak_footprint_bytes, footprint_);
        launch_rec rec;
        if (!entry_plans.empty()) {
            rec.pes = entry_plans;
        } else if (pe) {
            rec.pes.push_back(pe);
        }
        for (plan_entry *e : rec.pes) {
            ++e->inflight;
        }
        rec.reqs = group;
        rec.bytes = bytes;
        // the stream with the least work in flight (bytes of the launches queued on it; round
        // robin among equals), so a short request does not queue behind a long one when
        // another stream is free
        size_t si = next_stream_++ % streams_.size();
        for (size_t k = 0; k < streams_.size(); ++k) {
            const size_t c = (si + k) % streams_.size();
            if (stream_load_[c] < stream_load_[si]) {
                si = c;
            }
        }
        stream_load_[si] += bytes;
        rec.stream = si;
        hipStream_t stream = streams_[si];
        ++launching_;
        lk.unlock();
        // plan creation and the enqueue run without the lock (plan creation uploads tables)
        std::string msg;
        int status = SYCLFFT_HIP_STATUS_OK;
        std::unique_ptr<plan_entry> fresh;
        if (!pe) {
            syclfft_hip_plan_config_v1 cfg{};
            cfg.abi_version = SYCLFFT_DEVICE_PROVIDER_ABI_VERSION;
            cfg.scalar = head->r.scalar;
            cfg.direction = head->r.direction;
            cfg.rank = head->r.rank;
            for (uint32_t i = 0; i < 3; ++i) {
                cfg.lengths[i] = i < head->r.rank ? head->r.lengths[i] : 1;
            }
            cfg.batch_count = capacity;
            cfg.in_place = head->r.in_place;
            cfg.normalization = head->r.normalization;
            cfg.device = device_;
            char err[1024];
            void *h = nullptr;
            status = syclfft_hip_plan_create_ex(&cfg, &h, err, sizeof(err));
            if (status == SYCLFFT_HIP_STATUS_OK) {
                fresh = std::make_unique<plan_entry>();
                fresh->key = head->key;
                fresh->plan = h;
                fresh->capacity = capacity;
                // admitted on an upper bound (scratch for every transform); charged from here
                // on with what the plan holds, plus room for its pointer-array tables
                fresh->bytes = std::min(plan_bytes, sfh_plan_device_bytes(h) + (uint64_t(64) << 10));
                fresh->inflight = 1;
                pe = fresh.get();
            } else {
                msg = err;
            }
        }
        if (status == SYCLFFT_HIP_STATUS_OK && !entry_plans.empty()) {
            std::vector<void *> plans;
            std::vector<uint64_t> counts;
            std::vector<const void *> ins;
            std::vector<void *> outs;
            for (size_t k = 0; k < group.size(); ++k) {
                plans.push_back(entry_plans[k]->plan);
                counts.push_back(group[k]->r.batch_count);
                ins.push_back(group[k]->r.in);
                outs.push_back(group[k]->r.out);
            }
            char err[1024];
            rec.done = take_event();
            if (!rec.done) {
                status = SYCLFFT_HIP_STATUS_EXECUTION_FAILED;
                msg = "hipEventCreate failed";
            } else {
                status = sfh_execute_queue_checked_buffers(plans.data(), counts.data(), ins.data(), outs.data(),
                                                           plans.size(), stream, rec.done, err, sizeof(err));
                if (status != SYCLFFT_HIP_STATUS_OK) {
                    msg = err;
                    give_event(rec.done);
                    rec.done = nullptr;
                }
            }
        } else if (status == SYCLFFT_HIP_STATUS_OK) {
            std::vector<const void *> ins;
            std::vector<void *> outs;
            const uint64_t es = head->r.scalar == SYCLFFT_HIP_SCALAR_DOUBLE ? 16 : 8;
            const uint64_t tb = elems_of(head->r) * es;
            for (auto &q : group) {
                for (uint64_t b = 0; b < q->r.batch_count; ++b) {
                    ins.push_back(static_cast<const unsigned char *>(q->r.in) + b * tb);
                    outs.push_back(static_cast<unsigned char *>(q->r.out) + b * tb);
                }
            }
            char err[1024];
            rec.done = take_event();
            if (!rec.done) {
                status = SYCLFFT_HIP_STATUS_EXECUTION_FAILED;
                msg = "hipEventCreate failed";
            } else {
                status = sfh_plan_execute_many_checked_buffers(pe->plan, ins.size(), ins.data(), outs.data(), stream,
                                                               rec.done, err, sizeof(err));
                if (status != SYCLFFT_HIP_STATUS_OK) {
                    msg = err;
                    give_event(rec.done);
                    rec.done = nullptr;
                }
            }
        }
        const int64_t t = now_ns();
        lk.lock();
        --launching_;
        if (fresh) {
            footprint_ -= plan_bytes - fresh->bytes;
            by_key_[fresh->key].push_back(fresh.get());
            plans_.push_back(std::move(fresh));
            ++st_.plans_created;
        } else if (!pe) {
            footprint_ -= plan_bytes; // plan creation failed
        }
        if (rec.pes.empty() && pe) {
            rec.pes.push_back(pe); // the plan created for this launch (inflight counted at creation)
        }
        for (plan_entry *e : rec.pes) {
            e->last_use = t;
        }
        if (status != SYCLFFT_HIP_STATUS_OK) {
            footprint_ -= bytes;
            stream_load_[si] -= bytes;
            for (plan_entry *e : rec.pes) {
                --e->inflight;
            }
            fail(group, status, msg);
            cv_dispatch_.notify_one();
            return true;
        }
        for (auto &q : group) {
            q->t_launch = t;
            q->group = static_cast<uint32_t>(group.size());
        }
        inflight_.push_back(std::move(rec));
        ++st_.launches;
        cv_complete_.notify_one();
        return true;
    }

    void dispatch_loop()
    {
        (void)hipSetDevice(device_);
        std::unique_lock<std::mutex> lk(mu_);
        uint64_t seen = gen_;
        const auto ready = [&] {
            return (stop_ && pending_.empty()) || (!pending_.empty() && inflight_.size() < max_inflight_);
        };
        for (;;) {
            // stay awake briefly after work (a condition-variable wake-up costs tens of
            // microseconds of request latency), then sleep
            for (int spin = 0; spin < 2000 && !ready(); ++spin) {
                lk.unlock();
                std::this_thread::yield();
                lk.lock();
            }
            cv_dispatch_.wait(lk, ready);
            if (pending_.empty()) {
                if (stop_) {
                    return;
                }
                continue;
            }
            seen = gen_;
            if (!launch_next(lk)) {
                // wait until a completion changes the footprint
                cv_dispatch_.wait(lk, [&] { return stop_ || gen_ != seen; });
            }
        }
    }

    void complete_loop()
    {
        (void)hipSetDevice(device_);
        std::unique_lock<std::mutex> lk(mu_);
        for (;;) {
            if (inflight_.empty()) {
                if (stop_ && pending_.empty() && launching_ == 0) {
                    return;
                }
                cv_complete_.wait_for(lk, std::chrono::milliseconds(50));
                continue;
            }
            // query the launches' events without holding the lock (the dispatcher only
            // appends to inflight_ and only this thread erases, so iterators stay valid)
            std::vector<std::pair<std::list<launch_rec>::iterator, hipError_t>> ready;
            std::vector<std::list<launch_rec>::iterator> snap;
            for (auto it = inflight_.begin(); it != inflight_.end(); ++it) {
                snap.push_back(it);
            }
            lk.unlock();
            for (auto it : snap) {
                const hipError_t q = hipEventQuery(it->done);
                if (q != hipErrorNotReady) {
                    ready.emplace_back(it, q);
                }
            }
            lk.lock();
            bool any = false;
            for (auto &rq : ready) {
                auto it = rq.first;
                const hipError_t q = rq.second;
                const int64_t t = now_ns();
                for (auto &r : it->reqs) {
                    r->t_done = t;
                    if (q != hipSuccess) {
                        r->status = SYCLFFT_HIP_STATUS_EXECUTION_FAILED;
                        r->err = hipGetErrorString(q);
                        ++st_.failed;
                    } else {
                        ++st_.completed;
                    }
                    r->done = true;
                    --outstanding_;
                }
                footprint_ -= it->bytes;
                stream_load_[it->stream] -= it->bytes;
                for (plan_entry *e : it->pes) {
                    --e->inflight;
                    e->last_use = t;
                }
                give_event(it->done);
                inflight_.erase(it);
                any = true;
            }
            if (any) {
                ++gen_;
                cv_dispatch_.notify_one();
                cv_done_.notify_all();
            } else {
                lk.unlock();
                std::this_thread::yield();
                lk.lock();
            }
        }
    }

    int device_{};
    // merge requests of other shapes into work-queue launches (SYCLFFT_HIP_SERVE_QUEUE=0: off)
    bool queue_launches_{true};
    // size-first dispatch with an age bound (SYCLFFT_HIP_SERVE_SJF, _SJF_AGE_US)
    bool sjf_{false};
    int64_t sjf_age_ns_{500000};
    int64_t t0_{now_ns()};
    uint64_t budget_, max_batch_;
    size_t max_inflight_{};
    std::vector<hipStream_t> streams_;
    std::vector<uint64_t> stream_load_; // bytes of the launches in flight per stream
    size_t next_stream_{0};
    std::mutex mu_;
    std::condition_variable cv_dispatch_, cv_done_, cv_complete_;
    std::deque<std::shared_ptr<request>> pending_;
    std::unordered_map<uint64_t, std::shared_ptr<request>> table_;
    std::list<std::unique_ptr<plan_entry>> plans_;
    std::unordered_map<std::string, std::vector<plan_entry *>> by_key_; // resident plans per shape
    std::list<launch_rec> inflight_;
    uint64_t footprint_{0};
    uint64_t next_ticket_{1};
    uint64_t outstanding_{0};
    uint64_t gen_{0};
    int launching_{0};
    bool stop_{false};
    syclfft_hip_server_stats_v1 st_{};
    std::mutex ev_mu_;
    std::vector<hipEvent_t> events_;
    std::thread dispatcher_, completer_;
};

} // namespace sfh_srv

using sfh_srv::srv_error;

extern "C" void *syclfft_hip_server_create(
    int device, uint64_t memory_budget_bytes, uint32_t streams, uint64_t max_batch_transforms, char *err, size_t cap)
{
    try {
        sfh_srv::set_err(err, cap, "");
        return new sfh_srv::server(device, memory_budget_bytes, streams, max_batch_transforms);
    } catch (const std::exception &ex) {
        sfh_srv::set_err(err, cap, ex.what());
    } catch (...) {
        sfh_srv::set_err(err, cap, "unknown error");
    }
    return nullptr;
}

extern "C" int syclfft_hip_server_submit(
    void *server, const syclfft_hip_request_v1 *request, uint64_t *ticket, char *err, size_t cap)
{
    try {
        sfh_srv::set_err(err, cap, "");
        if (!server || !request || !ticket) {
            throw srv_error(SYCLFFT_HIP_STATUS_INVALID_ARGUMENT, "null server, request or ticket");
        }
        *ticket = static_cast<sfh_srv::server *>(server)->submit(*request);
        return SYCLFFT_HIP_STATUS_OK;
    } catch (const srv_error &ex) {
        sfh_srv::set_err(err, cap, ex.what());
        return ex.status;
    } catch (const std::exception &ex) {
        sfh_srv::set_err(err, cap, ex.what());
        return SYCLFFT_HIP_STATUS_EXECUTION_FAILED;
    }
}

extern "C" int syclfft_hip_server_wait(
    void *server, uint64_t ticket, syclfft_hip_completion_v1 *out, char *err, size_t cap)
{
    try {
        sfh_srv::set_err(err, cap, "");
        if (!server) {
            throw srv_error(SYCLFFT_HIP_STATUS_INVALID_ARGUMENT, "null server");
        }
        std::string msg;
        const int st = static_cast<sfh_srv::server *>(server)->wait(ticket, out, msg);
        sfh_srv::set_err(err, cap, msg);
        return st;
    } catch (const srv_error &ex) {
        sfh_srv::set_err(err, cap, ex.what());
        return ex.status;
    } catch (const std::exception &ex) {
        sfh_srv::set_err(err, cap, ex.what());
        return SYCLFFT_HIP_STATUS_EXECUTION_FAILED;
    }
}

extern "C" int syclfft_hip_server_drain(void *server, char *err, size_t cap)
{
    sfh_srv::set_err(err, cap, "");
    if (!server) {
        sfh_srv::set_err(err, cap, "null server");
        return SYCLFFT_HIP_STATUS_INVALID_ARGUMENT;
    }
    static_cast<sfh_srv::server *>(server)->drain();
    return SYCLFFT_HIP_STATUS_OK;
}

extern "C" void syclfft_hip_server_stats(void *server, syclfft_hip_server_stats_v1 *out)
{
    if (server && out) {
        static_cast<sfh_srv::server *>(server)->stats(out);
    }
}

extern "C" int syclfft_hip_server_replay(
    void *server, const syclfft_hip_request_v1 *requests, const double *arrival_seconds, size_t count,
    int sequential, syclfft_hip_completion_v1 *completions, char *err, size_t cap)
{
    sfh_srv::set_err(err, cap, "");
    if (!server || (count && (!requests || !completions))) {
        sfh_srv::set_err(err, cap, "null server, requests or completions");
        return SYCLFFT_HIP_STATUS_INVALID_ARGUMENT;
    }
    auto *srv = static_cast<sfh_srv::server *>(server);
    std::vector<uint64_t> tickets(count, 0);
    std::vector<int> status(count, SYCLFFT_HIP_STATUS_OK);
    std::vector<std::string> msgs(count);
    std::atomic<size_t> submitted{0};
    const auto submit_one = [&](size_t k) {
        try {
            tickets[k] = srv->submit(requests[k]);
        } catch (const srv_error &ex) {
            status[k] = ex.status;
            msgs[k] = ex.what();
        } catch (const std::exception &ex) {
            status[k] = SYCLFFT_HIP_STATUS_EXECUTION_FAILED;
            msgs[k] = ex.what();
        }
    };
    const auto wait_one = [&](size_t k) {
        completions[k] = syclfft_hip_completion_v1{};
        completions[k].tag = requests[k].tag;
        if (status[k] != SYCLFFT_HIP_STATUS_OK) {
            completions[k].status = status[k];
            return;
        }
        try {
            std::string m;
            status[k] = srv->wait(tickets[k], &completions[k], m);
            if (status[k] != SYCLFFT_HIP_STATUS_OK) {
                msgs[k] = m;
            }
        } catch (const srv_error &ex) {
            status[k] = ex.status;
            msgs[k] = ex.what();
        }
    };
    if (sequential) {
        for (size_t k = 0; k < count; ++k) {
            submit_one(k);
            wait_one(k);
        }
    } else {
        std::thread waiter([&] {
            for (size_t k = 0; k < count; ++k) {
                while (submitted.load(std::memory_order_acquire) <= k) {
                    std::this_thread::yield();
                }
                wait_one(k);
            }
        });
        const auto t0 = std::chrono::steady_clock::now();
        for (size_t k = 0; k < count; ++k) {
            if (arrival_seconds) {
                const auto due = t0 + std::chrono::duration_cast<std::chrono::steady_clock::duration>(
                                          std::chrono::duration<double>(arrival_seconds[k]));
                while (std::chrono::steady_clock::now() < due) {
                }
            }
            submit_one(k);
            submitted.store(k + 1, std::memory_order_release);
        }
        waiter.join();
    }
    for (size_t k = 0; k < count; ++k) {
        if (status[k] != SYCLFFT_HIP_STATUS_OK) {
            sfh_srv::set_err(err, cap, msgs[k]);
            return status[k];
        }
    }
    return SYCLFFT_HIP_STATUS_OK;
}

extern "C" void syclfft_hip_server_destroy(void *server)
{
    delete static_cast<sfh_srv::server *>(server);
}
