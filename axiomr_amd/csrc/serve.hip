// serve.hip - render server over the binning pipeline: requests (resident
// mesh + camera + resolution) arrive at any time, are admitted FIFO under an
// HBM budget for frame arenas and framebuffers, and are drawn in batches on
// several lanes (one stream and frame arena each) so that batches overlap
// on the device.  The scheduler itself is serve_core.hpp; this file is its
// device side and the C ABI.
//
// Each request is exactly one reference drawMesh (tiled_pipeline.cpp:
// 143-322) into its own framebuffer cleared to the server's clear colour
// and +inf depth (framebuffer.cpp:6-14), so its pixels do not depend on
// which other requests share its batch.
#include <algorithm>
#include <cstring>
#include <memory>
#include <new>
#include "internal.hpp"
#include "serve_core.hpp"

using namespace axr;
using namespace axr::serve;

namespace {

struct ServeReq {
    const axr_mesh* mesh;
    axr_draw_params params;
    int w, h;
};

constexpr uint64_t align256(uint64_t x) { return (x + 255) & ~255ull; }

// Framebuffer block layout: colour plane, depth plane, optional id plane,
// each 256-byte aligned.
uint64_t fb_plane(int w, int h) { return align256((uint64_t)w * h * 4); }
uint64_t fb_bytes(int w, int h, bool ids) { return fb_plane(w, h) * (ids ? 3 : 2); }

class HipDevice final : public Device<ServeReq> {
public:
    struct Lane {
        FrameArena fa;
        ViewDesc* d_views = nullptr;
        uint32_t* d_bases = nullptr;        // face, tile and pixel prefixes: 3 x (max_views + 1)
        uint32_t* d_tile_views = nullptr;
        ViewDesc* h_views = nullptr;        // pinned staging of the same
        uint32_t* h_bases = nullptr;
        uint32_t* h_tile_views = nullptr;
        hipEvent_t done = nullptr;
        KArgs ka{};
        FrameSize fs;
        int attempts = 0;
    };

    HipDevice(axr_ctx* ctx, const axr_server_config& cfg) : ctx_(ctx), cfg_(cfg) {
        tile_ = ctx->cfg.tile_size;
        wave_cap_ = (uint32_t)ctx->cfg.bin_capacity;
        raster_threads_ = ctx->cfg.raster_threads;
        small_footprint_ = ctx->cfg.small_footprint;
        mode_ = ctx->cfg.raster_mode;
        lanes_.resize(cfg.lanes);
    }

    ~HipDevice() override {
        (void)hipSetDevice(ctx_->device);
        for (Lane& l : lanes_) {
            arena_destroy(ctx_, l.fa);
            dev_free(ctx_, l.d_views, views_bytes());
            dev_free(ctx_, l.d_bases, bases_bytes());
            dev_free(ctx_, l.d_tile_views, tiles_bytes());
            if (l.h_views) (void)hipHostFree(l.h_views);
            if (l.h_bases) (void)hipHostFree(l.h_bases);
            if (l.h_tile_views) (void)hipHostFree(l.h_tile_views);
            if (l.done) (void)hipEventDestroy(l.done);
        }
    }

    int init() {
        (void)hipSetDevice(ctx_->device);
        for (Lane& l : lanes_) {
            int r = arena_create(ctx_, err_, l.fa);
            if (r) return r;
            // sized for a full batch up front (max_batch_faces faces over
            // max_batch_pixels pixels) so that serving does not grow it; only
            // a single request above those caps, which runs alone, grows it
            FrameSize fs;
            fs.faces = cfg_.max_batch_faces;
            fs.tiles = std::min<uint64_t>(cfg_.max_batch_tiles, lane_tiles());
            fs.pixels = mode_ == RASTER_VISIBILITY ? cfg_.max_batch_pixels : 0;
            fs.wave_cap = wave_cap_;
            if ((r = arena_ensure(ctx_, err_, l.fa, fs))) return r;
            if ((r = dev_alloc(ctx_, err_, (void**)&l.d_views, views_bytes())) ||
                (r = dev_alloc(ctx_, err_, (void**)&l.d_bases, bases_bytes())) ||
                (r = dev_alloc(ctx_, err_, (void**)&l.d_tile_views, tiles_bytes())))
                return r;
            hipError_t e = hipHostMalloc((void**)&l.h_views, views_bytes(), hipHostMallocDefault);
            if (e == hipSuccess) e = hipHostMalloc((void**)&l.h_bases, bases_bytes(), hipHostMallocDefault);
            if (e == hipSuccess) e = hipHostMalloc((void**)&l.h_tile_views, tiles_bytes(), hipHostMallocDefault);
            if (e == hipSuccess) e = hipEventCreateWithFlags(&l.done, hipEventDisableTiming);
            if (e != hipSuccess) return hip_fail(err_, e, "server lane");
        }
        return AXR_OK;
    }

    int lanes() const override { return (int)lanes_.size(); }
    void on_thread_start() override { (void)hipSetDevice(ctx_->device); }
    std::string error() const override { return err_.msg; }

    uint64_t fixed_bytes() const override {
        uint64_t b = 0;
        for (const Lane& l : lanes_) b += arena_bytes(l.fa) + views_bytes() + bases_bytes() + tiles_bytes();
        return b;
    }

    int alloc(uint64_t bytes, void** p) override { return dev_alloc(ctx_, err_, p, bytes); }
    void free(void* p, uint64_t bytes) override { dev_free(ctx_, p, bytes); }

    int launch(int lane, const std::vector<Job<ServeReq>*>& jobs) override {
        Lane& L = lanes_[lane];
        const uint32_t n = (uint32_t)jobs.size();
        const uint32_t stride = cfg_.max_batch_views + 1;
        uint64_t faces = 0, tiles = 0, pixels = 0;
        for (uint32_t i = 0; i < n; ++i) {
            const ServeReq& q = jobs[i]->pl;
            uint8_t* base = (uint8_t*)jobs[i]->fb.p;
            const uint64_t plane = fb_plane(q.w, q.h);
            ViewDesc v = make_view(q.mesh, (uint32_t*)base, (float*)(base + plane),
                                   cfg_.with_ids ? (uint32_t*)(base + 2 * plane) : nullptr, q.w, q.h, &q.params,
                                   tile_);
            v.face_base = (uint32_t)faces;
            v.tile_base = (uint32_t)tiles;
            v.pix_base = (uint32_t)pixels;
            L.h_views[i] = v;
            L.h_bases[i] = (uint32_t)faces;
            L.h_bases[stride + i] = (uint32_t)tiles;
            L.h_bases[2 * stride + i] = (uint32_t)pixels;
            faces += v.nfaces;
            tiles += (uint64_t)v.tiles_x * v.tiles_y;
            pixels += (uint64_t)q.w * q.h;
        }
        if (faces > MAX_BATCH_FACES || tiles > cfg_.max_batch_tiles || pixels > 0xFFFFFFFFull)
            return fail(err_, AXR_ERR_INVALID, "batch too large (%llu faces, %llu bins)", (unsigned long long)faces,
                        (unsigned long long)tiles);
        L.h_bases[n] = (uint32_t)faces;
        L.h_bases[stride + n] = (uint32_t)tiles;
        L.h_bases[2 * stride + n] = (uint32_t)pixels;
        KArgs& ka = L.ka;
        ka = KArgs{};
        ka.nviews = n;
        ka.total_faces = (uint32_t)faces;
        ka.total_tiles = (uint32_t)tiles;
        ka.tile_size = tile_;
        ka.raster_threads = raster_threads_;
        ka.small_footprint = small_footprint_;
        ka.mode = mode_;
        ka.total_pixels = (uint32_t)pixels;
        ka.shader = uniform_shader(L.h_views, n);
        ka.one = L.h_views[0];
        hipStream_t st = L.fa.stream;
        hipError_t e = hipSuccess;
        // every request starts from a cleared framebuffer; in visibility mode
        // the resolve pass writes that clear (k_resolve, ka.clear)
        const bool absorb = mode_ == RASTER_VISIBILITY;
        if (absorb) {
            ka.clear = 1;
            ka.clear_bgra = cfg_.clear_bgra;
            ka.clear_depth = __builtin_huge_valf();
        }
        if (n == 1) {
            // one request: the plain path, no tables
            const ViewDesc& v = L.h_views[0];
            if (!absorb)
                e = launch_clear(v.color, v.depth, v.ids, (size_t)v.W * v.H, cfg_.clear_bgra, __builtin_huge_valf(), st);
        } else {
            uint32_t t = 0;
            for (uint32_t i = 0; i < n; ++i)
                for (uint32_t k = L.h_bases[stride + i]; k < L.h_bases[stride + i + 1]; ++k) L.h_tile_views[t++] = i;
            e = hipMemcpyAsync(L.d_views, L.h_views, n * sizeof(ViewDesc), hipMemcpyHostToDevice, st);
            if (e == hipSuccess) e = hipMemcpyAsync(L.d_bases, L.h_bases, bases_bytes(), hipMemcpyHostToDevice, st);
            if (e == hipSuccess)
                e = hipMemcpyAsync(L.d_tile_views, L.h_tile_views, tiles * 4, hipMemcpyHostToDevice, st);
            ka.table = L.d_views;
            ka.face_bases = L.d_bases;
            ka.tile_bases = L.d_bases + stride;
            ka.tile_views = L.d_tile_views;
            if (e == hipSuccess && !absorb)
                e = launch_clear_views(L.d_views, L.d_bases + 2 * stride, n, (uint32_t)pixels, cfg_.clear_bgra,
                                       __builtin_huge_valf(), st);
        }
        if (e != hipSuccess) return hip_fail(err_, e, "server batch upload");
        L.fs = frame_size(ka, wave_cap_);
        L.attempts = 0;
        int r = arena_ensure(ctx_, err_, L.fa, L.fs);
        if (r) return r;
        return enqueue(L);
    }

    int query(int lane) override {
        Lane& L = lanes_[lane];
        hipError_t e = hipEventQuery(L.done);
        if (e == hipErrorNotReady) return 0;
        if (e != hipSuccess) return hip_fail(err_, e, "server batch");
        int r = arena_counters_after_wait(err_, L.fa);
        if (r) return r;
        if (!L.fa.host_counters[CNT_ERROR]) return 1;
        // arena overflow: the raster stage did not run, so the framebuffers
        // hold (at most) their clear; grow to the reported needs and redraw
        if (++L.attempts > 8) return fail(err_, AXR_ERR_OVERFLOW, "server batch kept overflowing its arena");
        r = arena_check_overflow(ctx_, err_, L.fa, L.fs);
        if (r != AXR_ERR_OVERFLOW) return r < 0 ? r : AXR_ERR_OVERFLOW;
        ++overflow_retries_;
        r = enqueue(L);
        return r ? r : 0;
    }

    uint64_t overflow_retries() const { return overflow_retries_; }
    int tile() const { return tile_; }

private:
    int enqueue(Lane& L) {
        int r = arena_launch(err_, L.fa, L.ka, nullptr);
        if (r) return r;
        // the counters come back only if the batch flags an error (query)
        const hipError_t e = hipEventRecord(L.done, L.fa.stream);
        return e == hipSuccess ? AXR_OK : hip_fail(err_, e, "server batch enqueue");
    }

    // Bins of a batch at the pixel cap: every pixel in bins of T x T, plus
    // the partial bins along two edges of each view.
    uint64_t lane_tiles() const {
        const uint64_t t2 = (uint64_t)tile_ * tile_;
        return cfg_.max_batch_pixels / t2 + (uint64_t)cfg_.max_batch_views * (2 * (MAX_FB_EDGE / tile_) + 1);
    }
    size_t views_bytes() const { return (size_t)cfg_.max_batch_views * sizeof(ViewDesc); }
    size_t bases_bytes() const { return (size_t)3 * (cfg_.max_batch_views + 1) * 4; }
    size_t tiles_bytes() const { return (size_t)cfg_.max_batch_tiles * 4; }

    axr_ctx* ctx_;
    axr_server_config cfg_;
    ErrSink err_;
    std::vector<Lane> lanes_;
    int tile_ = 16, raster_threads_ = 64, small_footprint_ = 12, mode_ = RASTER_BINNED;
    uint32_t wave_cap_ = 128;
    uint64_t overflow_retries_ = 0;
};

}  // namespace

struct axr_server {
    axr_ctx* ctx = nullptr;
    axr_server_config cfg{};
    std::unique_ptr<HipDevice> dev;
    std::unique_ptr<ServerCore<ServeReq>> core;
    std::string err;
};

namespace {

int server_fail(axr_server* s, int code, const std::string& msg) {
    s->err = msg;
    return code;
}

// Validates one request; fills the admission sizes.
int check_request(axr_server* s, const axr_serve_request* q, uint64_t* bytes, uint64_t* faces, uint64_t* tiles,
                  uint64_t* pixels) {
    if (!q || !q->mesh) return server_fail(s, AXR_ERR_INVALID, "request without a mesh");
    if (q->mesh->ctx != s->ctx) return server_fail(s, AXR_ERR_INVALID, "request mesh from another context");
    if (axr_check_view(q->mesh->nfaces_sub, q->width, q->height))
        return server_fail(s, AXR_ERR_INVALID, "request framebuffer size or face count out of range");
    if (!shader_ok(&q->params)) return server_fail(s, AXR_ERR_INVALID, "request names an unknown shader");
    const int t = s->dev->tile();
    *bytes = fb_bytes(q->width, q->height, s->cfg.with_ids != 0);
    *faces = q->mesh->nfaces_sub;
    *tiles = (uint64_t)((q->width + t - 1) / t) * ((q->height + t - 1) / t);
    *pixels = (uint64_t)q->width * (uint64_t)q->height;
    if (*tiles > s->cfg.max_batch_tiles)
        return server_fail(s, AXR_ERR_INVALID, "request has more bins than max_batch_tiles");
    return AXR_OK;
}

}  // namespace

extern "C" {

int axr_server_create(axr_ctx* ctx, const axr_server_config* cfg_in, axr_server** out) {
    if (!ctx || !out) return AXR_ERR_INVALID;
    *out = nullptr;
    axr_server_config cfg{};
    if (cfg_in) cfg = *cfg_in;
    if (cfg.lanes == 0) cfg.lanes = 3;
    if (cfg.max_batch_views == 0) cfg.max_batch_views = 64;
    if (cfg.max_batch_faces == 0) cfg.max_batch_faces = 8ull << 20;
    if (cfg.max_batch_tiles == 0) cfg.max_batch_tiles = 1ull << 20;
    if (cfg.max_batch_pixels == 0) cfg.max_batch_pixels = 32ull << 20;
    if (cfg.lanes < 1 || cfg.lanes > 8)
        return fail(ctx->err, AXR_ERR_INVALID, "server lanes must be in [1, 8] (got %d)", cfg.lanes);
    if (cfg.max_batch_views > 1024 || cfg.max_batch_faces > 0x40000000ull || cfg.max_batch_tiles > 0x10000000ull ||
        cfg.max_batch_pixels > 0xFFFFFFFFull)
        return fail(ctx->err, AXR_ERR_INVALID, "server batch limits too large");
    (void)hipSetDevice(ctx->device);
    axr_server* s = new (std::nothrow) axr_server();
    if (!s) return AXR_ERR_OOM;
    s->ctx = ctx;
    s->cfg = cfg;
    s->dev.reset(new (std::nothrow) HipDevice(ctx, cfg));
    if (!s->dev) { delete s; return AXR_ERR_OOM; }
    int r = s->dev->init();
    if (r) {
        ctx->err.msg = "server: " + s->dev->error();
        delete s;
        return r;
    }
    if (cfg.hbm_budget && s->dev->fixed_bytes() > cfg.hbm_budget) {
        fail(ctx->err, AXR_ERR_BUDGET, "server frame arenas (%llu bytes) exceed its HBM budget (%llu)",
             (unsigned long long)s->dev->fixed_bytes(), (unsigned long long)cfg.hbm_budget);
        delete s;
        return AXR_ERR_BUDGET;
    }
    CoreConfig cc;
    cc.max_views = cfg.max_batch_views;
    cc.max_faces = cfg.max_batch_faces;
    cc.max_tiles = cfg.max_batch_tiles;
    cc.max_pixels = cfg.max_batch_pixels;
    cc.budget = cfg.hbm_budget;
    s->core.reset(new (std::nothrow) ServerCore<ServeReq>(cc, s->dev.get()));
    if (!s->core) { delete s; return AXR_ERR_OOM; }
    s->core->start();
    *out = s;
    return AXR_OK;
}

int axr_server_destroy(axr_server* s) {
    if (!s) return AXR_ERR_INVALID;
    (void)hipSetDevice(s->ctx->device);
    if (s->core) {
        s->core->stop();
        s->core->release_pool();
    }
    s->core.reset();
    s->dev.reset();
    delete s;
    return AXR_OK;
}

const char* axr_server_last_error(axr_server* s) {
    if (!s) return "null server";
    std::string core = s->core ? s->core->last_error() : std::string();
    if (!core.empty() && s->err.empty()) s->err = core;
    return s->err.c_str();
}

int axr_server_submit(axr_server* s, const axr_serve_request* q, uint64_t* ticket) {
    if (!s || !ticket) return AXR_ERR_INVALID;
    uint64_t bytes, faces, tiles, pixels;
    int r = check_request(s, q, &bytes, &faces, &tiles, &pixels);
    if (r) return r;
    ServeReq pl{q->mesh, q->params, q->width, q->height};
    r = s->core->submit(pl, bytes, faces, tiles, pixels, now_s(), false, nullptr, ticket);
    if (r) return server_fail(s, r, s->core->last_error());
    return AXR_OK;
}

int axr_server_poll(axr_server* s, uint64_t ticket, int* state) {
    if (!s || !state) return AXR_ERR_INVALID;
    int status = 0;
    const int st = s->core->state(ticket, &status);
    if (st < 0) return server_fail(s, AXR_ERR_INVALID, "unknown or released ticket");
    *state = st;
    return st == JOB_FAILED ? status : AXR_OK;
}

int axr_server_wait(axr_server* s, uint64_t ticket, double timeout_s) {
    if (!s) return AXR_ERR_INVALID;
    const int r = s->core->wait(ticket, timeout_s);
    if (r == AXR_ERR_INVALID) return server_fail(s, r, "unknown or released ticket");
    if (r < 0 && r != AXR_ERR_NOT_READY) return server_fail(s, r, s->dev->error());
    return r;
}

int axr_server_read(axr_server* s, uint64_t ticket, uint8_t* bgra, float* depth, uint32_t* ids) {
    if (!s) return AXR_ERR_INVALID;
    Block b;
    ServeReq q;
    const int r = s->core->done_block(ticket, &b, &q);
    if (r) return server_fail(s, r, r == AXR_ERR_INVALID ? "unknown or released ticket" : "request not done");
    if (ids && !s->cfg.with_ids) return server_fail(s, AXR_ERR_INVALID, "server keeps no id planes");
    (void)hipSetDevice(s->ctx->device);
    const size_t n = (size_t)q.w * q.h;
    const uint64_t plane = fb_plane(q.w, q.h);
    const uint8_t* base = (const uint8_t*)b.p;
    // the request's lane has completed: plain copies, no stream ordering needed
    hipError_t e = hipSuccess;
    if (bgra) e = hipMemcpy(bgra, base, n * 4, hipMemcpyDeviceToHost);
    if (e == hipSuccess && depth) e = hipMemcpy(depth, base + plane, n * 4, hipMemcpyDeviceToHost);
    if (e == hipSuccess && ids) e = hipMemcpy(ids, base + 2 * plane, n * 4, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return server_fail(s, AXR_ERR_HIP, std::string("server read: ") + hipGetErrorString(e));
    return AXR_OK;
}

int axr_server_release(axr_server* s, uint64_t ticket) {
    if (!s) return AXR_ERR_INVALID;
    const int r = s->core->release(ticket);
    if (r == AXR_ERR_INVALID) return server_fail(s, r, "unknown or released ticket");
    if (r) return server_fail(s, r, "request still queued or running");
    return AXR_OK;
}

int axr_server_pause(axr_server* s, int pause) {
    if (!s) return AXR_ERR_INVALID;
    s->core->pause(pause != 0);
    return AXR_OK;
}

int axr_server_get_stats(axr_server* s, axr_server_stats* out) {
    if (!s || !out) return AXR_ERR_INVALID;
    const CoreStats c = s->core->stats();
    *out = axr_server_stats{};
    out->submitted = c.submitted;
    out->completed = c.completed;
    out->failed = c.failed;
    out->rejected = c.rejected;
    out->batches = c.batches;
    out->batched_views = c.batched_views;
    out->max_views_in_batch = c.max_views_in_batch;
    out->launches_overlapped = c.launches_overlapped;
    out->max_lanes_busy = c.max_lanes_busy;
    out->deferred = c.deferred;
    out->fb_reserved_bytes = c.fb_reserved;
    out->fb_peak_bytes = c.fb_peak;
    out->fb_in_use_bytes = c.fb_in_use;
    out->fixed_bytes = c.fixed;
    out->queue_peak = c.queue_peak;
    out->overflow_retries = s->dev->overflow_retries();
    return AXR_OK;
}

int axr_server_replay(axr_server* s, uint32_t n, const axr_serve_request* reqs, const double* arrival_s,
                      double* latency_s) {
    if (!s || (n && (!reqs || !arrival_s || !latency_s))) return AXR_ERR_INVALID;
    std::vector<ServeReq> pls(n);
    std::vector<uint64_t> bytes(n), faces(n), tiles(n), pixels(n);
    for (uint32_t i = 0; i < n; ++i) {
        int r = check_request(s, &reqs[i], &bytes[i], &faces[i], &tiles[i], &pixels[i]);
        if (r) return r;
        if (i && arrival_s[i] < arrival_s[i - 1])
            return server_fail(s, AXR_ERR_INVALID, "replay arrival times must be non-decreasing");
        pls[i] = ServeReq{reqs[i].mesh, reqs[i].params, reqs[i].width, reqs[i].height};
    }
    const int r = s->core->replay(n, pls.data(), bytes.data(), faces.data(), tiles.data(), pixels.data(), arrival_s,
                                  latency_s);
    if (r) return server_fail(s, r, s->core->last_error());
    return AXR_OK;
}

}  // extern "C"
