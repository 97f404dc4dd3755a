// Host-only tests of the render server's scheduler and admission controller
// (axiomr_amd/csrc/serve_core.hpp) against a fake device whose batches
// complete after a set time.  Built and run by tests/test_serve_core.py
// under AddressSanitizer + UBSan and under ThreadSanitizer.
#include <cassert>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <set>
#include "../../axiomr_amd/csrc/serve_core.hpp"

using namespace axr::serve;

#define CHECK(c)                                                                  \
    do {                                                                          \
        if (!(c)) {                                                               \
            std::fprintf(stderr, "%s:%d: CHECK failed: %s\n", __FILE__, __LINE__, #c); \
            std::exit(1);                                                         \
        }                                                                         \
    } while (0)

struct Req {
    int id = -1;
    bool fail_launch = false;
    bool fail_run = false;
};

struct FakeDevice : Device<Req> {
    int nlanes;
    uint64_t fixed;
    double batch_s;                          // simulated duration of one batch
    struct LaneState {
        bool busy = false;
        double t_end = 0;
        bool fail = false;
    };
    std::vector<LaneState> ls;
    std::vector<int> launch_order;           // request ids in launch order
    std::vector<std::vector<int>> batches;
    uint64_t allocs = 0, frees = 0, live_bytes = 0;
    std::set<void*> live;
    uintptr_t next_addr = 0x10000;

    FakeDevice(int lanes, uint64_t fixed_bytes, double batch_seconds)
        : nlanes(lanes), fixed(fixed_bytes), batch_s(batch_seconds), ls(lanes) {}
    int lanes() const override { return nlanes; }
    uint64_t fixed_bytes() const override { return fixed; }
    int alloc(uint64_t bytes, void** p) override {
        *p = (void*)next_addr;
        next_addr += bytes;
        live.insert(*p);
        live_bytes += bytes;
        ++allocs;
        return S_OK;
    }
    void free(void* p, uint64_t bytes) override {
        CHECK(live.erase(p) == 1);
        live_bytes -= bytes;
        ++frees;
    }
    int launch(int lane, const std::vector<Job<Req>*>& jobs) override {
        CHECK(!ls[lane].busy);
        std::vector<int> ids;
        bool fail = false;
        for (auto* j : jobs) {
            CHECK(j->fb.p != nullptr && j->fb.bytes >= j->fb_bytes);
            CHECK(live.count(j->fb.p) == 1);
            if (j->pl.fail_launch) return -3;
            fail |= j->pl.fail_run;
            ids.push_back(j->pl.id);
            launch_order.push_back(j->pl.id);
        }
        batches.push_back(ids);
        ls[lane] = LaneState{true, now_s() + batch_s, fail};
        return S_OK;
    }
    int query(int lane) override {
        CHECK(ls[lane].busy);
        if (now_s() < ls[lane].t_end) return 0;
        ls[lane].busy = false;
        return ls[lane].fail ? -3 : 1;
    }
};

static const uint64_t MB = 1ull << 20;

static void test_fifo_and_batching() {
    FakeDevice dev(3, 100 * MB, 0.002);
    CoreConfig cc;
    cc.max_views = 4;
    cc.max_faces = 1000;
    ServerCore<Req> core(cc, &dev);
    core.start();
    std::vector<uint64_t> t(60);
    for (int i = 0; i < 60; ++i) {
        Req r;
        r.id = i;
        CHECK(core.submit(r, 2 * MB, 300, 100, 0, now_s(), false, nullptr, &t[i]) == S_OK);
    }
    for (int i = 0; i < 60; ++i) CHECK(core.wait(t[i], 10.0) == S_OK);
    core.stop();
    for (int i = 0; i < 60; ++i) CHECK(dev.launch_order[i] == i);   // FIFO
    for (auto& b : dev.batches) CHECK(b.size() <= 3);                   // 3 x 300 faces <= 1000 < 4 x 300
    CoreStats s = core.stats();
    CHECK(s.completed == 60 && s.failed == 0);
    CHECK(s.max_lanes_busy == 3);
    CHECK(s.launches_overlapped > 0);
    CHECK(s.max_views_in_batch >= 2);
    for (int i = 0; i < 60; ++i) CHECK(core.release(t[i]) == S_OK);
    CHECK(core.release(t[0]) == S_INVALID);        // stale ticket
    core.release_pool();
    CHECK(dev.live.empty());
    std::puts("fifo_and_batching ok");
}

static void test_oversized_single_request_runs_alone() {
    FakeDevice dev(2, 0, 0.0005);
    CoreConfig cc;
    cc.max_faces = 1000;
    ServerCore<Req> core(cc, &dev);
    core.start();
    uint64_t a, b, c;
    Req r;
    r.id = 0;
    CHECK(core.submit(r, MB, 5000, 1, 0, now_s(), false, nullptr, &a) == S_OK);
    r.id = 1;
    CHECK(core.submit(r, MB, 10, 1, 0, now_s(), false, nullptr, &b) == S_OK);
    r.id = 2;
    CHECK(core.submit(r, MB, 10, 1, 0, now_s(), false, nullptr, &c) == S_OK);
    CHECK(core.wait(a, 10) == S_OK && core.wait(b, 10) == S_OK && core.wait(c, 10) == S_OK);
    core.stop();
    for (auto& bt : dev.batches)
        for (int id : bt)
            if (id == 0) CHECK(bt.size() == 1);
    core.release_pool();
    std::puts("oversized_single ok");
}

static void test_pixel_cap() {
    // batches hold at most max_pixels framebuffer pixels; a single larger
    // request runs alone
    FakeDevice dev(1, 0, 0.0005);
    CoreConfig cc;
    cc.max_views = 16;
    cc.max_pixels = 1000;
    ServerCore<Req> core(cc, &dev);
    core.pause(true);   // queue everything first, so batching is deterministic
    core.start();
    std::vector<uint64_t> t(9);
    const uint64_t px[9] = {400, 400, 400, 5000, 100, 100, 1000, 1, 999};
    for (int i = 0; i < 9; ++i) {
        Req r;
        r.id = i;
        CHECK(core.submit(r, MB, 1, 1, px[i], now_s(), false, nullptr, &t[i]) == S_OK);
    }
    core.pause(false);
    for (int i = 0; i < 9; ++i) CHECK(core.wait(t[i], 10.0) == S_OK);
    core.stop();
    const std::vector<std::vector<int>> want = {{0, 1}, {2}, {3}, {4, 5}, {6}, {7, 8}};
    CHECK(dev.batches == want);
    for (int i = 0; i < 9; ++i) CHECK(core.release(t[i]) == S_OK);
    core.release_pool();
    std::puts("pixel_cap ok");
}

static void test_budget_defers_and_rejects() {
    // budget: fixed 10 MB + room for two 4 MB-class blocks
    const uint64_t blk = class_bytes(class_of(4 * MB));
    FakeDevice dev(3, 10 * MB, 0.001);
    CoreConfig cc;
    cc.budget = 10 * MB + 2 * blk;
    ServerCore<Req> core(cc, &dev);
    core.start();
    uint64_t big;
    Req r;
    CHECK(core.submit(r, 3 * blk, 1, 1, 0, now_s(), false, nullptr, &big) == S_BUDGET);   // can never fit
    std::vector<uint64_t> t(12);
    for (int i = 0; i < 12; ++i) {
        r.id = i;
        CHECK(core.submit(r, 4 * MB, 1, 1, 0, now_s(), false, nullptr, &t[i]) == S_OK);
    }
    // the client releases results as they complete; the rest wait in the queue
    for (int i = 0; i < 12; ++i) {
        CHECK(core.wait(t[i], 10.0) == S_OK);
        if (i + 1 < 12) CHECK(core.state(t[i + 1]) != JOB_FREE);
        CHECK(core.release(t[i]) == S_OK);
    }
    core.stop();
    CoreStats s = core.stats();
    CHECK(s.rejected == 1 && s.completed == 12);
    CHECK(s.fb_peak <= 2 * blk);
    CHECK(s.deferred > 0);
    CHECK(dev.allocs <= 2);                      // blocks are reused, not reallocated
    for (int i = 0; i < 12; ++i) CHECK(dev.launch_order[i] == i);
    core.release_pool();
    CHECK(dev.live.empty() && dev.live_bytes == 0);
    std::puts("budget ok");
}

static void test_eviction_of_idle_blocks() {
    // room for one block of the large class only: a small idle block must be
    // returned to the device before the large request can be admitted
    const uint64_t small = class_bytes(class_of(2 * MB)), large = class_bytes(class_of(16 * MB));
    FakeDevice dev(1, 0, 0.0);
    CoreConfig cc;
    cc.budget = large;
    ServerCore<Req> core(cc, &dev);
    Req r;
    uint64_t a, b;
    CHECK(core.submit(r, 2 * MB, 1, 1, 0, now_s(), false, nullptr, &a) == S_OK);
    while (core.state(a) != JOB_DONE) core.step();
    CHECK(core.release(a) == S_OK);
    CHECK(core.submit(r, 16 * MB, 1, 1, 0, now_s(), false, nullptr, &b) == S_OK);
    for (int k = 0; k < 100 && core.state(b) != JOB_DONE; ++k) core.step();
    CHECK(core.state(b) == JOB_DONE);
    CHECK(dev.frees == 1 && dev.live_bytes == large);
    (void)small;
    core.release_pool();
    std::puts("eviction ok");
}

static void test_failures() {
    FakeDevice dev(2, 0, 0.0005);
    ServerCore<Req> core(CoreConfig{}, &dev);
    core.start();
    std::vector<uint64_t> t(20);
    for (int i = 0; i < 20; ++i) {
        Req r;
        r.id = i;
        r.fail_launch = (i == 5);
        r.fail_run = (i == 11);
        CHECK(core.submit(r, MB, 1, 1, 0, now_s(), false, nullptr, &t[i]) == S_OK);
    }
    int failed = 0;
    for (int i = 0; i < 20; ++i) {
        const int w = core.wait(t[i], 10.0);
        if (w != S_OK) {
            ++failed;
            CHECK(w == -3);
            int status = 0;
            CHECK(core.state(t[i], &status) == JOB_FAILED && status == -3);
        }
    }
    core.stop();
    CHECK(failed >= 2);     // at least the two planted requests (their batch mates fail with them)
    CoreStats s = core.stats();
    CHECK(s.completed + s.failed == 20);
    for (int i = 0; i < 20; ++i) CHECK(core.release(t[i]) == S_OK);
    core.release_pool();
    CHECK(dev.live.empty());
    std::puts("failures ok");
}

static void test_release_rules() {
    FakeDevice dev(1, 0, 0.0);
    ServerCore<Req> core(CoreConfig{}, &dev);
    Req r;
    uint64_t a;
    CHECK(core.submit(r, MB, 1, 1, 0, now_s(), false, nullptr, &a) == S_OK);
    CHECK(core.release(a) == S_NOT_READY);            // still queued
    CHECK(core.state(a) == JOB_QUEUED);
    core.step();
    while (core.state(a) != JOB_DONE) core.step();
    Block b;
    Req back;
    CHECK(core.done_block(a, &b, &back) == S_OK && b.p != nullptr);
    CHECK(core.release(a) == S_OK);
    CHECK(core.state(a) == S_INVALID);
    CHECK(core.done_block(a, &b, &back) == S_INVALID);
    uint64_t a2;
    CHECK(core.submit(r, MB, 1, 1, 0, now_s(), false, nullptr, &a2) == S_OK);
    CHECK((a2 & 0xFFFFFFFFu) == (a & 0xFFFFFFFFu) && a2 != a);   // slot reused, new generation
    CHECK(core.state(a) == S_INVALID);
    core.release_pool();
    std::puts("release_rules ok");
}

static void test_pause_batches_a_burst() {
    FakeDevice dev(2, 0, 0.0);
    CoreConfig cc;
    cc.max_views = 4;
    ServerCore<Req> core(cc, &dev);
    core.pause(true);
    std::vector<uint64_t> t(10);
    for (int i = 0; i < 10; ++i) {
        Req r;
        r.id = i;
        CHECK(core.submit(r, MB, 1, 1, 0, now_s(), false, nullptr, &t[i]) == S_OK);
    }
    CHECK(!core.step() && core.queued() == 10 && dev.batches.empty());   // nothing admitted while paused
    core.pause(false);
    core.step();
    CHECK(dev.batches.size() == 2 && dev.batches[0].size() == 4 && dev.batches[1].size() == 4);
    for (int k = 0; k < 100 && core.queued(); ++k) core.step();
    while (core.state(t[9]) != JOB_DONE) core.step();
    CHECK(dev.batches.size() == 3 && dev.batches[2].size() == 2);
    CoreStats s = core.stats();
    CHECK(s.max_lanes_busy == 2 && s.launches_overlapped == 1);
    core.release_pool();
    std::puts("pause ok");
}

static void test_replay_open_loop() {
    FakeDevice dev(3, 0, 0.0003);
    ServerCore<Req> core(CoreConfig{}, &dev);
    core.start();
    const uint32_t n = 400;
    std::vector<Req> pls(n);
    std::vector<uint64_t> bytes(n, 2 * MB), faces(n, 10), tiles(n, 10);
    std::vector<double> arr(n), lat(n);
    for (uint32_t i = 0; i < n; ++i) {
        pls[i].id = (int)i;
        arr[i] = i * 0.0001;       // 10k requests per second: batches form behind busy lanes
        if (i % 7 == 0) bytes[i] = 9 * MB;
    }
    CHECK(core.replay(n, pls.data(), bytes.data(), faces.data(), tiles.data(), nullptr, arr.data(), lat.data()) == S_OK);
    core.stop();
    for (uint32_t i = 0; i < n; ++i) CHECK(lat[i] >= 0.0 && lat[i] < 5.0);
    CoreStats s = core.stats();
    CHECK(s.completed == n && s.fb_in_use == 0);
    CHECK(s.batches < n);               // continuous batching merged requests
    core.release_pool();
    CHECK(dev.live.empty());
    std::puts("replay ok");
}

static void test_concurrent_clients() {
    FakeDevice dev(4, 0, 0.0002);
    ServerCore<Req> core(CoreConfig{}, &dev);
    core.start();
    std::vector<std::thread> th;
    std::atomic<int> ok{0};
    for (int c = 0; c < 4; ++c)
        th.emplace_back([&, c] {
            for (int i = 0; i < 50; ++i) {
                Req r;
                r.id = c * 1000 + i;
                uint64_t t;
                if (core.submit(r, (1 + (i % 3)) * MB, 5, 5, 0, now_s(), false, nullptr, &t) != S_OK) continue;
                if (core.wait(t, 10.0) == S_OK && core.release(t) == S_OK) ++ok;
            }
        });
    for (auto& t : th) t.join();
    core.stop();
    CHECK(ok.load() == 200);
    core.release_pool();
    CHECK(dev.live.empty());
    std::puts("concurrent_clients ok");
}

int main() {
    CHECK(class_of(class_bytes(5)) == 5 && class_of(class_bytes(5) + 1) == 6);
    for (int c = 1; c < kClasses; ++c) CHECK(class_bytes(c) > class_bytes(c - 1));
    CHECK(class_bytes(class_of(16588800)) < 16588800ull * 6 / 5);   // 1080p colour+depth: < 20% slack
    test_fifo_and_batching();
    test_oversized_single_request_runs_alone();
    test_pixel_cap();
    test_budget_defers_and_rejects();
    test_eviction_of_idle_blocks();
    test_failures();
    test_release_rules();
    test_pause_batches_a_burst();
    test_replay_open_loop();
    test_concurrent_clients();
    std::puts("ALL OK");
    return 0;
}
