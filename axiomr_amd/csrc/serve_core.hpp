// serve_core.hpp - the host side of the render server: request table,
// FIFO admission under an HBM budget, continuous batching onto lanes, and
// latency bookkeeping.  Pure C++17 with no HIP: the device is reached
// through the Device interface (serve.hip implements it with streams,
// events and the binning kernels; the CPU tests implement it with a fake
// clock), so the admission controller and scheduler are tested on the host,
// under the sanitizers.
//
// The reference has no serving layer; the per-request work is one drawMesh
// (tiled_pipeline.cpp:143-322) into a cleared framebuffer of the requested
// size.
#pragma once
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdint>
#include <deque>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

namespace axr {
namespace serve {

inline double now_s() {
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}

enum JobState : int { JOB_FREE = 0, JOB_QUEUED = 1, JOB_RUNNING = 2, JOB_DONE = 3, JOB_FAILED = 4 };

// Return codes, numerically equal to the C ABI's axr_status values.
enum : int { S_OK = 0, S_INVALID = -1, S_OOM = -2, S_BUDGET = -5, S_NOT_READY = -7 };

// Framebuffer block of the size-class pool.
struct Block {
    void* p = nullptr;
    uint64_t bytes = 0;   // class size (what the pool reserved)
    int cls = -1;
};

// Size classes: 64 KiB * 2^(c/4), rounded up to 4 KiB (strictly
// increasing); at most 19% of a block is slack.
constexpr int kClasses = 96;
inline uint64_t class_bytes(int c) {
    const double b = 65536.0 * std::pow(2.0, c / 4.0);
    return ((uint64_t)std::ceil(b / 4096.0)) * 4096ull;
}
inline int class_of(uint64_t bytes) {
    for (int c = 0; c < kClasses; ++c)
        if (class_bytes(c) >= bytes) return c;
    return -1;
}

template <class Payload>
struct Job {
    Payload pl{};
    uint64_t fb_bytes = 0, faces = 0, tiles = 0, pixels = 0;
    Block fb;
    int state = JOB_FREE;
    int status = S_OK;
    uint32_t gen = 0;
    bool auto_release = false;
    double t_arrival = 0, t_launch = 0, t_done = 0;
    double* out_latency = nullptr;   // replay: written at completion
    uint32_t* out_left = nullptr;    // replay: decremented at completion
};

// What the scheduler needs from the device.  All calls come from the
// scheduler thread, except release_all at shutdown.
template <class Payload>
struct Device {
    virtual ~Device() = default;
    virtual int lanes() const = 0;
    virtual void on_thread_start() {}     // first call on the scheduler thread
    virtual std::string error() const { return ""; }
    // bytes held outside the framebuffer pool (frame arenas, batch tables)
    virtual uint64_t fixed_bytes() const = 0;
    virtual int alloc(uint64_t bytes, void** p) = 0;   // S_OK or an error code
    virtual void free(void* p, uint64_t bytes) = 0;
    // Enqueue one batch on `lane`; S_OK or an error code (nothing enqueued).
    virtual int launch(int lane, const std::vector<Job<Payload>*>& jobs) = 0;
    // 0 while the lane's batch runs, 1 when it completed, < 0 on failure.
    // A device that retries internally (arena overflow) keeps returning 0.
    virtual int query(int lane) = 0;
};

struct CoreConfig {
    uint32_t max_views = 64;
    uint64_t max_faces = 8ull << 20;
    uint64_t max_tiles = 1ull << 20;
    uint64_t max_pixels = 0;   // framebuffer pixels per batch; 0 = unlimited
    uint64_t budget = 0;   // framebuffers + fixed bytes; 0 = unlimited
};

struct CoreStats {
    uint64_t submitted = 0, completed = 0, failed = 0, rejected = 0;
    uint64_t batches = 0, batched_views = 0, max_views_in_batch = 0;
    uint64_t launches_overlapped = 0;   // launched while another lane was busy
    uint32_t max_lanes_busy = 0;
    uint64_t deferred = 0;              // admission stopped for lack of budget
    uint64_t fb_reserved = 0, fb_peak = 0, fb_in_use = 0, fixed = 0;
    uint64_t queue_peak = 0;
};

template <class Payload>
class ServerCore {
public:
    using JobT = Job<Payload>;

    ServerCore(const CoreConfig& cfg, Device<Payload>* dev) : cfg_(cfg), dev_(dev) {
        lanes_.resize(dev->lanes());
    }
    ~ServerCore() { stop(); }

    void start() {
        std::lock_guard<std::mutex> lk(mu_);
        if (running_) return;
        running_ = true;
        thread_ = std::thread([this] { run(); });
    }

    // Stops the scheduler thread after the lanes drain.  Queued requests
    // stay queued (they fail with S_NOT_READY if waited on).
    void stop() {
        {
            std::lock_guard<std::mutex> lk(mu_);
            if (!running_) return;
            running_ = false;
        }
        cv_.notify_all();
        if (thread_.joinable()) thread_.join();
        std::unique_lock<std::mutex> lk(mu_);
        while (busy_lanes() > 0) {
            harvest();
            if (busy_lanes() > 0) {
                lk.unlock();
                std::this_thread::yield();
                lk.lock();
            }
        }
    }

    // Returns every pool block to the device (after stop()).
    void release_pool() {
        std::lock_guard<std::mutex> lk(mu_);
        for (auto& j : jobs_)
            if (j.fb.p) { dev_->free(j.fb.p, j.fb.bytes); stats_.fb_reserved -= j.fb.bytes; j.fb = Block{}; }
        for (int c = 0; c < kClasses; ++c) {
            for (void* p : free_[c]) { dev_->free(p, class_bytes(c)); stats_.fb_reserved -= class_bytes(c); }
            free_[c].clear();
        }
    }

    int submit(const Payload& pl, uint64_t fb_bytes, uint64_t faces, uint64_t tiles, uint64_t pixels, double arrival,
               bool auto_release, double* out_latency, uint64_t* ticket, uint32_t* out_left = nullptr) {
        const int cls = class_of(fb_bytes);
        std::unique_lock<std::mutex> lk(mu_);
        if (cls < 0 || (cfg_.budget && dev_->fixed_bytes() + class_bytes(cls) > cfg_.budget)) {
            ++stats_.rejected;
            err_ = "request framebuffer of " + std::to_string(fb_bytes) + " bytes can never fit the HBM budget";
            return S_BUDGET;
        }
        uint32_t slot;
        if (!free_slots_.empty()) {
            slot = free_slots_.back();
            free_slots_.pop_back();
        } else {
            if (jobs_.size() >= 0xFFFFFFFFull) return S_OOM;
            slot = (uint32_t)jobs_.size();
            jobs_.emplace_back();
        }
        JobT& j = jobs_[slot];
        const uint32_t gen = j.gen + 1;
        j = JobT{};
        j.gen = gen;
        j.pl = pl;
        j.fb_bytes = fb_bytes;
        j.faces = faces;
        j.tiles = tiles;
        j.pixels = pixels;
        j.state = JOB_QUEUED;
        j.auto_release = auto_release;
        j.t_arrival = arrival;
        j.out_latency = out_latency;
        j.out_left = out_left;
        queue_.push_back(slot);
        ++stats_.submitted;
        stats_.queue_peak = std::max<uint64_t>(stats_.queue_peak, queue_.size());
        if (ticket) *ticket = ((uint64_t)gen << 32) | slot;
        lk.unlock();
        cv_.notify_all();
        return S_OK;
    }

    // state of a ticket (JobState) or S_INVALID for an unknown/released one
    int state(uint64_t ticket, int* status = nullptr) {
        std::lock_guard<std::mutex> lk(mu_);
        JobT* j = find(ticket);
        if (!j) return S_INVALID;
        if (status) *status = j->status;
        return j->state;
    }

    // Waits until the request is done or failed; S_NOT_READY on timeout.
    int wait(uint64_t ticket, double timeout_s) {
        std::unique_lock<std::mutex> lk(mu_);
        const auto deadline = std::chrono::steady_clock::now() + std::chrono::duration<double>(timeout_s);
        for (;;) {
            JobT* j = find(ticket);
            if (!j) return S_INVALID;
            if (j->state == JOB_DONE) return S_OK;
            if (j->state == JOB_FAILED) return j->status;
            if (!running_) return S_NOT_READY;
            if (done_cv_.wait_until(lk, deadline) == std::cv_status::timeout) {
                j = find(ticket);
                if (j && (j->state == JOB_DONE || j->state == JOB_FAILED)) continue;
                return S_NOT_READY;
            }
        }
    }

    // The framebuffer block and payload of a completed request.  The block
    // stays valid until the request is released.
    int done_block(uint64_t ticket, Block* fb, Payload* pl) {
        std::lock_guard<std::mutex> lk(mu_);
        JobT* j = find(ticket);
        if (!j) return S_INVALID;
        if (j->state != JOB_DONE) return j->state == JOB_FAILED ? j->status : S_NOT_READY;
        if (fb) *fb = j->fb;
        if (pl) *pl = j->pl;
        return S_OK;
    }

    int release(uint64_t ticket) {
        std::lock_guard<std::mutex> lk(mu_);
        JobT* j = find(ticket);
        if (!j) return S_INVALID;
        if (j->state == JOB_RUNNING || j->state == JOB_QUEUED) return S_NOT_READY;
        free_job((uint32_t)(ticket & 0xFFFFFFFFu));
        return S_OK;
    }

    CoreStats stats() {
        std::lock_guard<std::mutex> lk(mu_);
        CoreStats s = stats_;
        s.fixed = dev_->fixed_bytes();
        return s;
    }

    // While paused, requests queue but none is admitted; resuming admits
    // the queue in order (so a burst submitted while paused is batched).
    void pause(bool on) {
        {
            std::lock_guard<std::mutex> lk(mu_);
            paused_ = on;
        }
        cv_.notify_all();
    }

    std::string last_error() {
        std::lock_guard<std::mutex> lk(mu_);
        return err_;
    }

    // Open-loop driver: request i is submitted at start + arrival_s[i] (the
    // calling thread spins on the clock), released at completion, and its
    // latency (completion - scheduled arrival) is written to latency_s[i]
    // (negative for a failed request).  Returns when all are done.
    // pixels may be null (no pixel sizes: the max_pixels cap does not apply).
    int replay(uint32_t n, const Payload* pls, const uint64_t* fb_bytes, const uint64_t* faces,
               const uint64_t* tiles, const uint64_t* pixels, const double* arrival_s, double* latency_s) {
        for (uint32_t i = 0; i < n; ++i) latency_s[i] = -1.0;
        uint32_t left = n;   // guarded by mu_
        const double t0 = now_s();
        int rc = S_OK;
        for (uint32_t i = 0; i < n; ++i) {
            const double t = t0 + arrival_s[i];
            while (now_s() < t) std::this_thread::yield();
            int r = submit(pls[i], fb_bytes[i], faces[i], tiles[i], pixels ? pixels[i] : 0, t, true, &latency_s[i],
                           nullptr, &left);
            if (r != S_OK) {
                rc = r;
                std::lock_guard<std::mutex> lk(mu_);
                --left;
            }
        }
        std::unique_lock<std::mutex> lk(mu_);
        done_cv_.wait(lk, [&] { return left == 0 || !running_; });
        return left == 0 ? rc : S_NOT_READY;
    }

    // One scheduler iteration (harvest finished lanes, then admit onto free
    // lanes).  The scheduler thread loops on it; single-threaded tests call
    // it directly.  Returns true if anything changed.
    bool step() {
        std::lock_guard<std::mutex> lk(mu_);
        bool changed = harvest();
        changed |= admit();
        return changed;
    }

    size_t queued() {
        std::lock_guard<std::mutex> lk(mu_);
        return queue_.size();
    }

private:
    struct Lane {
        bool busy = false;
        std::vector<uint32_t> slots;
    };

    JobT* find(uint64_t ticket) {
        const uint32_t slot = (uint32_t)(ticket & 0xFFFFFFFFu), gen = (uint32_t)(ticket >> 32);
        if (slot >= jobs_.size()) return nullptr;
        JobT& j = jobs_[slot];
        return (j.state == JOB_FREE || j.gen != gen) ? nullptr : &j;
    }

    void free_job(uint32_t slot) {
        JobT& j = jobs_[slot];
        if (j.fb.p) {
            free_[j.fb.cls].push_back(j.fb.p);
            stats_.fb_in_use -= j.fb.bytes;
            j.fb = Block{};
        }
        j.state = JOB_FREE;
        free_slots_.push_back(slot);
    }

    int busy_lanes() const {
        int n = 0;
        for (const Lane& l : lanes_) n += l.busy ? 1 : 0;
        return n;
    }

    // A framebuffer block for `bytes`, from the free list, a new device
    // allocation within the budget, or after returning idle blocks of other
    // classes to the device.  False: defer.
    bool take_block(uint64_t bytes, Block* out) {
        const int c = class_of(bytes);
        const uint64_t cb = class_bytes(c);
        if (!free_[c].empty()) {
            *out = Block{free_[c].back(), cb, c};
            free_[c].pop_back();
            stats_.fb_in_use += cb;
            return true;
        }
        auto fits = [&] { return !cfg_.budget || dev_->fixed_bytes() + stats_.fb_reserved + cb <= cfg_.budget; };
        for (int o = kClasses - 1; o >= 0 && !fits(); --o) {
            while (!free_[o].empty() && !fits()) {
                dev_->free(free_[o].back(), class_bytes(o));
                free_[o].pop_back();
                stats_.fb_reserved -= class_bytes(o);
            }
        }
        if (!fits()) return false;
        void* p = nullptr;
        if (dev_->alloc(cb, &p) != S_OK) return false;
        stats_.fb_reserved += cb;
        stats_.fb_peak = std::max(stats_.fb_peak, stats_.fb_reserved);
        stats_.fb_in_use += cb;
        *out = Block{p, cb, c};
        return true;
    }

    bool harvest() {
        bool changed = false;
        for (int l = 0; l < (int)lanes_.size(); ++l) {
            Lane& lane = lanes_[l];
            if (!lane.busy) continue;
            const int r = dev_->query(l);
            if (r == 0) continue;
            const double t = now_s();
            for (uint32_t slot : lane.slots) finish(slot, r > 0 ? S_OK : r, t);
            lane.slots.clear();
            lane.busy = false;
            changed = true;
        }
        if (changed) done_cv_.notify_all();
        return changed;
    }

    void finish(uint32_t slot, int status, double t) {
        JobT& j = jobs_[slot];
        j.t_done = t;
        j.status = status;
        j.state = status == S_OK ? JOB_DONE : JOB_FAILED;
        if (status == S_OK) ++stats_.completed; else ++stats_.failed;
        if (j.out_latency) *j.out_latency = status == S_OK ? t - j.t_arrival : -1.0;
        if (j.out_left) --*j.out_left;
        if (j.auto_release) free_job(slot);
    }

    bool admit() {
        if (paused_) return false;
        bool changed = false;
        std::vector<JobT*> batch;
        for (int l = 0; l < (int)lanes_.size() && !queue_.empty(); ++l) {
            Lane& lane = lanes_[l];
            if (lane.busy) continue;
            batch.clear();
            uint64_t faces = 0, tiles = 0, pixels = 0;
            bool deferred = false;
            while (!queue_.empty() && batch.size() < cfg_.max_views) {
                JobT& j = jobs_[queue_.front()];
                // a single request above a cap runs alone
                if (!batch.empty() && (faces + j.faces > cfg_.max_faces || tiles + j.tiles > cfg_.max_tiles ||
                                       (cfg_.max_pixels && pixels + j.pixels > cfg_.max_pixels)))
                    break;
                if (!take_block(j.fb_bytes, &j.fb)) { deferred = true; break; }
                faces += j.faces;
                tiles += j.tiles;
                pixels += j.pixels;
                lane.slots.push_back(queue_.front());
                batch.push_back(&j);
                queue_.pop_front();
            }
            if (deferred) ++stats_.deferred;
            if (batch.empty()) break;   // the head does not fit yet: FIFO, nothing may pass it
            const int others = busy_lanes();
            const double t = now_s();
            for (JobT* j : batch) {
                j->state = JOB_RUNNING;
                j->t_launch = t;
            }
            const int r = dev_->launch(l, batch);
            if (r != S_OK) {
                err_ = "batch launch failed: " + dev_->error();
                for (uint32_t slot : lane.slots) finish(slot, r, t);
                lane.slots.clear();
                done_cv_.notify_all();
                changed = true;
                continue;
            }
            lane.busy = true;
            ++stats_.batches;
            stats_.batched_views += batch.size();
            stats_.max_views_in_batch = std::max<uint64_t>(stats_.max_views_in_batch, batch.size());
            if (others > 0) ++stats_.launches_overlapped;
            stats_.max_lanes_busy = std::max<uint32_t>(stats_.max_lanes_busy, (uint32_t)(others + 1));
            changed = true;
            if (deferred) break;
        }
        return changed;
    }

    void run() {
        dev_->on_thread_start();
        std::unique_lock<std::mutex> lk(mu_);
        while (running_) {
            bool changed = harvest();
            changed |= admit();
            if (changed) continue;
            if (busy_lanes() == 0) {
                // idle: sleep until a submission (or a release that may
                // un-defer the queue head) or shutdown
                cv_.wait_for(lk, std::chrono::milliseconds(queue_.empty() ? 50 : 1));
            } else {
                lk.unlock();
                std::this_thread::yield();
                lk.lock();
            }
        }
    }

    CoreConfig cfg_;
    Device<Payload>* dev_;
    std::mutex mu_;
    std::condition_variable cv_, done_cv_;
    std::thread thread_;
    bool running_ = false;
    bool paused_ = false;
    std::vector<JobT> jobs_;
    std::vector<uint32_t> free_slots_;
    std::deque<uint32_t> queue_;
    std::vector<Lane> lanes_;
    std::vector<void*> free_[kClasses];
    CoreStats stats_;
    std::string err_;
};

}  // namespace serve
}  // namespace axr
