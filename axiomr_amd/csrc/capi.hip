// capi.hip - host side of the C ABI declared in include/axiomr.h: context,
// device resources, the frame arena, drawMesh, the graph-captured batch.
#include <algorithm>
#include <atomic>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include "internal.hpp"

using namespace axr;

struct axr_batch {
    axr_ctx* ctx;
    uint32_t nviews;
    std::vector<ViewDesc> views;
    std::vector<axr_fb*> fbs;      // the views' framebuffers (for their pending clears)
    std::vector<const axr_mesh*> meshes;
    bool holds_refs = false;       // fbs and meshes count this batch in batch_refs
    ViewDesc* d_views = nullptr;
    uint32_t* d_face_bases = nullptr;
    uint32_t* d_tile_bases = nullptr;
    uint32_t* d_pix_bases = nullptr;
    uint32_t* d_tile_views = nullptr;
    uint32_t total_pixels = 0;
    // the vertex stage once per vertex of each mesh the views share (k_vary):
    // one job per (mesh, shader kind, the uniforms the vertex stage reads)
    std::vector<VaryJob> vary_jobs;
    std::vector<size_t> vary_bytes;     // the output buffer of each job
    VaryJob* d_vary_jobs = nullptr;
    uint32_t* d_vary_firsts = nullptr;
    bool clear = false;            // clear every view's framebuffer before drawing
    uint32_t clear_bgra = 0;
    float clear_depth = 0.0f;
    FrameSize fs;                  // arena needs, with the tunables the batch was built with
    KArgs ka{};
    size_t bytes = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    uint64_t captured_gen = ~0ull;
    uint64_t clear_version = 0, captured_clear = ~0ull;
};

namespace axr {

int fail(ErrSink& err, int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    err.msg = buf;
    return code;
}

int hip_fail(ErrSink& err, hipError_t e, const char* what) {
    return fail(err, AXR_ERR_HIP, "%s: %s", what, hipGetErrorString(e));
}

int dev_alloc(axr_ctx* ctx, ErrSink& err, void** p, size_t bytes) {
    *p = nullptr;
    if (bytes == 0) return AXR_OK;
    {
        std::lock_guard<std::mutex> lk(ctx->mu);
        if (ctx->budget && ctx->in_use + ctx->arena_bytes + bytes > ctx->budget)
            return fail(err, AXR_ERR_BUDGET, "allocation of %zu bytes exceeds the HBM budget (%llu in use of %llu)",
                        bytes, (unsigned long long)(ctx->in_use + ctx->arena_bytes), (unsigned long long)ctx->budget);
        ctx->in_use += bytes;
    }
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) {
        *p = nullptr;
        std::lock_guard<std::mutex> lk(ctx->mu);
        ctx->in_use -= bytes;
        return fail(err, AXR_ERR_OOM, "hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
    }
    return AXR_OK;
}

void dev_free(axr_ctx* ctx, void* p, size_t bytes) {
    if (!p) return;
    (void)hipFree(p);
    std::lock_guard<std::mutex> lk(ctx->mu);
    ctx->in_use -= bytes;
}

int upload(ErrSink& err, hipStream_t st, void* dst, const void* src, size_t bytes) {
    if (!bytes) return AXR_OK;
    hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    return e == hipSuccess ? AXR_OK : hip_fail(err, e, "upload");
}

static void free_buf(axr_ctx* ctx, DevBuf& b) {
    if (b.p) (void)hipFree(b.p);
    {
        std::lock_guard<std::mutex> lk(ctx->mu);
        ctx->arena_bytes -= b.bytes;
    }
    b.p = nullptr;
    b.bytes = 0;
}

static int grow(axr_ctx* ctx, ErrSink& err, FrameArena& fa, DevBuf& b, size_t bytes) {
    if (b.bytes >= bytes) return AXR_OK;
    bytes = std::max(bytes, b.bytes + b.bytes / 4);
    hipError_t e = hipStreamSynchronize(fa.stream);
    if (e != hipSuccess) return hip_fail(err, e, "sync before arena growth");
    free_buf(ctx, b);
    {
        std::lock_guard<std::mutex> lk(ctx->mu);
        if (ctx->budget && ctx->in_use + ctx->arena_bytes + bytes > ctx->budget)
            return fail(err, AXR_ERR_BUDGET, "frame arena of %zu more bytes exceeds the HBM budget", bytes);
        ctx->arena_bytes += bytes;
    }
    e = hipMalloc(&b.p, bytes);
    if (e != hipSuccess) {
        b.p = nullptr;
        std::lock_guard<std::mutex> lk(ctx->mu);
        ctx->arena_bytes -= bytes;
        return fail(err, AXR_ERR_OOM, "arena hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
    }
    b.bytes = bytes;
    ++fa.gen;
    return AXR_OK;
}

FrameSize frame_size(const KArgs& ka, uint32_t wave_cap) {
    FrameSize fs;
    fs.faces = ka.total_faces;
    fs.tiles = ka.total_tiles;
    fs.pixels = ka.mode == RASTER_VISIBILITY ? ka.total_pixels : 0;
    fs.wave_cap = wave_cap;
    return fs;
}

int arena_ensure(axr_ctx* ctx, ErrSink& err, FrameArena& fa, const FrameSize& fs, const ArenaNeeds& need) {
    const uint64_t faces = fs.faces, tiles = fs.tiles;
    const uint32_t wave_cap = fs.wave_cap;
    if (faces > 0x7FFFFFFFull || tiles > 0x7FFFFFFFull || fs.pixels > 0xFFFFFFFFull)
        return fail(err, AXR_ERR_INVALID, "draw too large (%llu faces, %llu bins, %llu pixels)",
                    (unsigned long long)faces, (unsigned long long)tiles, (unsigned long long)fs.pixels);
    Arena& a = fa.a;
    // visibility mode (fs.pixels != 0) keeps no triangle records per face
    // slot (only the first clip record of a clipped face, 4 bytes) and no
    // bin pairs; binned mode keeps a 48-byte record per face slot
    const bool vis = fs.pixels != 0;
    uint64_t clip_faces = std::max<uint64_t>({4096, faces / 16, need.clip_faces, a.cap_clip_faces});
    uint64_t clip_recs = std::max<uint64_t>({2 * clip_faces, need.clip_recs, a.cap_clip_recs});
    // one pair region per setup wave (setup runs 4 waves per 256 faces)
    uint64_t waves = vis ? a.cap_waves : std::max<uint64_t>(4 * ((faces + 255) / 256), a.cap_waves);
    uint64_t ovf = vis ? a.cap_ovf : std::max<uint64_t>({faces / 4 + 16 * tiles + 4096, need.pairs, a.cap_ovf});
    uint64_t nbins = vis ? a.cap_bins : std::max<uint64_t>({2 * faces + 16 * tiles + 4096, need.pairs, a.cap_bins});
    uint64_t scan_blocks = (tiles + 4095) / 4096 + 1;
    // visibility mode: large triangles and their 256-pixel work items
    uint64_t big_recs = 0, big_items = 0;
    if (fs.pixels) {
        big_recs = std::max<uint64_t>({4096, faces / 8, need.big_recs, a.cap_big_recs});
        big_items = std::max<uint64_t>({16384, faces / 4 + fs.pixels / 32, need.big_items, a.cap_big_items});
    }
    const uint64_t gen_start = fa.gen;
    int r;
    if ((r = grow(ctx, err, fa, fa.counters, CNT_COUNT * 4)) ||
        (r = grow(ctx, err, fa, fa.rec, std::max<uint64_t>(faces, 1) * (vis ? 4 : sizeof(TriRec)))) ||
        (r = grow(ctx, err, fa, fa.clip_faces, clip_faces * 4)) ||
        (r = grow(ctx, err, fa, fa.clip_rec, clip_recs * sizeof(TriRec))) ||
        (r = grow(ctx, err, fa, fa.clip_verts, clip_recs * 3 * sizeof(Vertex56))) ||
        (r = grow(ctx, err, fa, fa.tile_count, (tiles + 1) * 4)) ||
        (r = grow(ctx, err, fa, fa.tile_off, (tiles + 1) * 4)) ||
        (r = grow(ctx, err, fa, fa.pairs, (waves * wave_cap + ovf) * 16)) ||
        (r = grow(ctx, err, fa, fa.wave_pairs, waves * 4)) ||
        (r = grow(ctx, err, fa, fa.bins, nbins * 4)) ||
        (r = grow(ctx, err, fa, fa.active, (tiles + 1) * 4)) ||
        (r = grow(ctx, err, fa, fa.scan_sums, scan_blocks * 8)) ||
        (r = grow(ctx, err, fa, fa.vkeys, fs.pixels * 8)) ||
        (r = grow(ctx, err, fa, fa.big_rec, big_recs * sizeof(TriRec))) ||
        (r = grow(ctx, err, fa, fa.big_start, big_recs * 4)) ||
        (r = grow(ctx, err, fa, fa.vcol, fs.color_slots * 4)))
        return r;
    a.counters = (uint32_t*)fa.counters.p;
    a.host_err = fa.host_err;
    a.rec = (TriRec*)fa.rec.p;
    a.clip_base = (uint32_t*)fa.rec.p;
    a.clip_faces = (uint32_t*)fa.clip_faces.p;
    a.clip_rec = (TriRec*)fa.clip_rec.p;
    a.clip_verts = (Vertex56*)fa.clip_verts.p;
    a.tile_count = (uint32_t*)fa.tile_count.p;
    a.tile_off = (uint32_t*)fa.tile_off.p;
    a.pairs = (uint4*)fa.pairs.p;
    a.wave_pairs = (uint32_t*)fa.wave_pairs.p;
    a.bins = (uint32_t*)fa.bins.p;
    a.active = (uint32_t*)fa.active.p;
    a.scan_sums = (unsigned long long*)fa.scan_sums.p;
    a.vkeys = (unsigned long long*)fa.vkeys.p;
    a.big_rec = (TriRec*)fa.big_rec.p;
    a.big_start = (uint32_t*)fa.big_start.p;
    a.vcol = (uint32_t*)fa.vcol.p;
    a.cap_clip_faces = (uint32_t)std::min<uint64_t>(fa.clip_faces.bytes / 4, 0xFFFFFFFFull);
    a.cap_clip_recs = (uint32_t)std::min<uint64_t>(
        {fa.clip_rec.bytes / sizeof(TriRec), fa.clip_verts.bytes / (3 * sizeof(Vertex56)), 0x7FFFFFFFull});
    if (a.wave_cap != wave_cap) ++fa.gen;   // the pair layout changed: captured graphs are stale
    a.wave_cap = wave_cap;
    a.cap_waves = (uint32_t)std::min<uint64_t>(
        {fa.wave_pairs.bytes / 4, 0x7FFFFFFFull / wave_cap, fa.pairs.bytes / 16 / wave_cap});
    a.cap_ovf = (uint32_t)std::min<uint64_t>(fa.pairs.bytes / 16 - (uint64_t)a.cap_waves * wave_cap, 0x7FFFFFFFull);
    a.cap_bins = (uint32_t)std::min<uint64_t>(fa.bins.bytes / 4, 0x7FFFFFFFull);
    a.cap_vkeys = (uint32_t)std::min<uint64_t>(fa.vkeys.bytes / 8, 0xFFFFFFFFull);
    a.cap_big_recs = (uint32_t)std::min<uint64_t>(
        {fa.big_rec.bytes / sizeof(TriRec), fa.big_start.bytes / 4, 0x7FFFFFFFull});
    // items are only counted (no storage): the cap bounds the work of k_big
    a.cap_big_items = (uint32_t)std::min<uint64_t>(std::max<uint64_t>(big_items, a.cap_big_items), 0x7FFFFFFFull);
    if (fa.gen != gen_start) fa.dirty = true;
    return AXR_OK;
}

// Zero the whole counter block (including the sticky error and the
// high-water marks) and mark the bin counts for a k_prep.  Not used inside
// captured graphs.
static int arena_reset(ErrSink& err, FrameArena& fa) {
    std::memset(fa.host_counters, 0, CNT_COUNT * 4);
    *fa.host_err = 0;
    hipError_t e = hipMemcpyAsync(fa.counters.p, fa.host_counters, CNT_COUNT * 4, hipMemcpyHostToDevice, fa.stream);
    if (e == hipSuccess) e = hipStreamSynchronize(fa.stream);
    fa.dirty = true;
    return e == hipSuccess ? AXR_OK : hip_fail(err, e, "reset counters");
}

// Signal memory for the framebuffer handoff between arena streams: one
// 8-byte signal per frame arena, waited on with hipStreamWaitValue64 and
// written with hipStreamWriteValue64.  Signals are never freed (the runtime
// serves them from memory the host and the command processor share);
// an arena that is destroyed returns its signal to a per-device free list
// for the next one.  An arena zeroes the signal it takes (nothing waits on
// it then); the values it writes afterwards come from one process-wide
// counter, so they only grow and a wait for "at least v" is met by the
// write of v and by nothing written before it.
namespace {
std::mutex g_sig_mu;
std::vector<std::pair<int, uint64_t*>> g_sig_free;   // (device, signal)
std::atomic<uint64_t> g_sig_value{0};

// AXR_STREAM_WAIT_VALUE=0: hand over with events instead.
bool sig_enabled() {
    static const bool on = [] {
        const char* v = std::getenv("AXR_STREAM_WAIT_VALUE");
        return !(v && v[0] == '0');
    }();
    return on;
}

uint64_t* sig_acquire(int device) {
    if (!sig_enabled()) return nullptr;
    std::lock_guard<std::mutex> lk(g_sig_mu);
    for (size_t i = 0; i < g_sig_free.size(); ++i)
        if (g_sig_free[i].first == device) {
            uint64_t* p = g_sig_free[i].second;
            g_sig_free.erase(g_sig_free.begin() + (long)i);
            return p;
        }
    int can = 0;
    if (hipDeviceGetAttribute(&can, hipDeviceAttributeCanUseStreamWaitValue, device) != hipSuccess || !can)
        return nullptr;
    void* p = nullptr;
    if (hipExtMallocWithFlags(&p, 8, hipMallocSignalMemory) != hipSuccess) return nullptr;
    return (uint64_t*)p;
}

void sig_release(int device, uint64_t* p) {
    if (!p) return;
    std::lock_guard<std::mutex> lk(g_sig_mu);
    g_sig_free.emplace_back(device, p);
}
}  // namespace

// The pinned block of an arena: the counters as last copied back, then the
// error word the kernels set, 128-byte aligned past them.
constexpr uint32_t HOST_ERR_WORD = (CNT_COUNT + 31) / 32 * 32 + 32;

int arena_create(axr_ctx* ctx, ErrSink& err, FrameArena& fa) {
    hipError_t e = hipStreamCreateWithFlags(&fa.stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipHostMalloc((void**)&fa.host_counters, (HOST_ERR_WORD + 32) * 4, hipHostMallocDefault);
    if (e != hipSuccess) return hip_fail(err, e, "frame arena stream");
    fa.host_err = fa.host_counters + HOST_ERR_WORD;
    *fa.host_err = 0;
    FrameSize fs;
    fs.faces = 1024;
    fs.tiles = 1024;
    fs.wave_cap = (uint32_t)ctx->cfg.bin_capacity;
    int r = arena_ensure(ctx, err, fa, fs);
    // null: this arena's draws hand over with events.  The value starts at
    // 0 (arena_reset waits for the stream).
    fa.sig = sig_acquire(ctx->device);
    if (fa.sig && hipStreamWriteValue64(fa.stream, fa.sig, 0, 0) != hipSuccess) {
        sig_release(ctx->device, fa.sig);
        fa.sig = nullptr;
    }
    return r ? r : arena_reset(err, fa);
}

constexpr int kArenaBufs = 16;
static void arena_buffers(FrameArena& fa, DevBuf* (&out)[kArenaBufs]) {
    DevBuf* bufs[kArenaBufs] = {&fa.counters,  &fa.rec,      &fa.clip_faces, &fa.clip_rec,   &fa.clip_verts, &fa.tile_count,
                                &fa.tile_off,  &fa.pairs,    &fa.wave_pairs, &fa.bins,       &fa.active,     &fa.scan_sums,
                                &fa.vkeys,     &fa.big_rec,  &fa.big_start,  &fa.vcol};
    for (int i = 0; i < kArenaBufs; ++i) out[i] = bufs[i];
}

uint64_t arena_bytes(const FrameArena& fa) {
    DevBuf* bufs[kArenaBufs];
    arena_buffers(const_cast<FrameArena&>(fa), bufs);
    uint64_t b = 0;
    for (DevBuf* d : bufs) b += d->bytes;
    return b;
}

void arena_destroy(axr_ctx* ctx, FrameArena& fa) {
    if (fa.stream) (void)hipStreamSynchronize(fa.stream);
    DevBuf* bufs[kArenaBufs];
    arena_buffers(fa, bufs);
    for (DevBuf* b : bufs) free_buf(ctx, *b);
    if (fa.host_counters) (void)hipHostFree(fa.host_counters);
    if (fa.stream) (void)hipStreamDestroy(fa.stream);
    sig_release(ctx->device, fa.sig);
    fa = FrameArena{};
}

int arena_read_counters(ErrSink& err, FrameArena& fa) {
    hipError_t e = hipMemcpyAsync(fa.host_counters, fa.counters.p, CNT_COUNT * 4, hipMemcpyDeviceToHost, fa.stream);
    if (e == hipSuccess) e = hipStreamSynchronize(fa.stream);
    return e == hipSuccess ? AXR_OK : hip_fail(err, e, "read counters");
}

// Every device error bit also sets the pinned flag, so without the flag the
// device counters hold no error.
int arena_counters_after_wait(ErrSink& err, FrameArena& fa) {
    if (__atomic_load_n(fa.host_err, __ATOMIC_ACQUIRE) == 0) {
        fa.host_counters[CNT_ERROR] = 0;
        return AXR_OK;
    }
    return arena_read_counters(err, fa);
}

static int arena_prep_if_dirty(ErrSink& err, FrameArena& fa) {
    if (!fa.dirty) return AXR_OK;
    hipError_t e = launch_prep(fa.a, (uint32_t)(fa.tile_count.bytes / 4), fa.stream);
    if (e != hipSuccess) return hip_fail(err, e, "arena reset");
    fa.dirty = false;
    return AXR_OK;
}

int arena_launch(ErrSink& err, FrameArena& fa, const KArgs& ka, StageEvents* ev, const FbWait& wait) {
    int r = arena_prep_if_dirty(err, fa);
    if (r) return r;
    hipError_t e = launch_frame(ka, fa.a, fa.stream, ev, wait);
    return e == hipSuccess ? AXR_OK : hip_fail(err, e, "kernel launch");
}

int arena_check_overflow(axr_ctx* ctx, ErrSink& err, FrameArena& fa, const FrameSize& fs) {
    const uint32_t* c = fa.host_counters;
    if (!c[CNT_ERROR]) return AXR_OK;
    // a pass the host judged empty had work: never skip passes again here
    if (c[CNT_ERROR] & ERR_SKIPPED) {
        ctx->no_pass_skips = true;
        ++ctx->stats.skip_faults;
    }
    auto grown = [](uint64_t n, uint64_t num, uint64_t den) { return n ? n + n * num / den : 0; };
    ArenaNeeds need;
    need.clip_faces = grown(c[CNT_NEED_CLIP_FACES], 1, 2);
    need.clip_recs = grown(c[CNT_NEED_CLIP_RECS], 1, 2);
    need.pairs = grown(c[CNT_NEED_PAIRS], 1, 4);
    need.big_recs = grown(c[CNT_NEED_BIG_RECS], 1, 2);
    need.big_items = grown(c[CNT_NEED_BIG_ITEMS], 1, 2);
    // every arena of the context is grown to at least this on its next draw
    ArenaNeeds& fl = ctx->need_floor;
    fl.clip_faces = std::max(fl.clip_faces, need.clip_faces);
    fl.clip_recs = std::max(fl.clip_recs, need.clip_recs);
    fl.pairs = std::max(fl.pairs, need.pairs);
    fl.big_recs = std::max(fl.big_recs, need.big_recs);
    fl.big_items = std::max(fl.big_items, need.big_items);
    const uint32_t seen[5] = {c[CNT_NEED_CLIP_FACES], c[CNT_NEED_CLIP_RECS], c[CNT_NEED_PAIRS], c[CNT_NEED_BIG_RECS],
                              c[CNT_NEED_BIG_ITEMS]};   // the reset below clears the counters
    int r = arena_ensure(ctx, err, fa, fs, need);
    if (r) return r;
    if ((r = arena_reset(err, fa))) return r;
    return fail(err, AXR_ERR_OVERFLOW, "frame arena overflow (clip faces %u, clip records %u, bin pairs %u, "
                "large triangles %u, large-triangle items %u); arena grown, frames since the last sync were not drawn",
                seen[0], seen[1], seen[2], seen[3], seen[4]);
}

static FrameArena& arena_no(axr_ctx* ctx, int k) { return k ? ctx->alt[k - 1] : ctx->main; }

// Tools that collect hardware counters serialize kernel dispatches across
// queues, and work spread over streams that wait on each other can then
// stall; under them (or with AXR_OVERLAP=0) every draw goes to the
// context's own stream, one frame after the other.
static bool overlap_enabled() {
    static const bool on = [] {
        const char* v = std::getenv("AXR_OVERLAP");
        if (v && v[0] == '0') return false;
        for (const char* k : {"ROCPROF_COUNTER_COLLECTION", "ROCPROF_COUNTERS", "HSA_TOOLS_LIB", "ROCP_TOOL_LIB"})
            if (std::getenv(k)) return false;
        return true;
    }();
    return on;
}

int ctx_join(axr_ctx* ctx) {
    hipError_t e = hipSuccess;
    for (int k = 1; k < axr_ctx::kArenas && e == hipSuccess; ++k) {
        if (!ctx->alt[k - 1].stream || !ctx->alt_unjoined[k - 1]) continue;
        e = hipEventRecord(ctx->join_ev[k], ctx->alt[k - 1].stream);
        if (e == hipSuccess) e = hipStreamWaitEvent(ctx->main.stream, ctx->join_ev[k], 0);
        if (e == hipSuccess) ctx->alt_unjoined[k - 1] = false;
    }
    if (e != hipSuccess) return hip_fail(ctx->err, e, "stream order");
    ctx->last_arena = 0;
    ctx->last_value = 0;
    return AXR_OK;
}

int ctx_sync_all(axr_ctx* ctx) {
    hipError_t e = hipStreamSynchronize(ctx->main.stream);
    for (int k = 1; k < axr_ctx::kArenas && e == hipSuccess; ++k)
        if (ctx->alt[k - 1].stream) e = hipStreamSynchronize(ctx->alt[k - 1].stream);
    return e == hipSuccess ? AXR_OK : hip_fail(ctx->err, e, "hipStreamSynchronize");
}

int fb_settle(axr_fb* fb) {
    if (!fb->clear_pending) return AXR_OK;
    int r = ctx_join(fb->ctx);
    if (r) return r;
    hipError_t e = launch_clear(fb->color, fb->depth, fb->ids, (size_t)fb->w * fb->h, fb->clear_bgra, fb->clear_depth,
                                fb->ctx->main.stream);
    if (e != hipSuccess) return hip_fail(fb->ctx->err, e, "framebuffer clear");
    fb->clear_pending = false;
    return AXR_OK;
}

PassSkips pass_skips(const axr_mesh* m, const mat4& mvp, int W, int H, int footprint) {
    PassSkips s;
    if (!std::isfinite(m->radius)) return s;
    // row k of the clip transform: clip_k = r[k] . (p, 1)
    double r[4][4];
    for (int k = 0; k < 4; ++k) {
        const float* e = &mvp.c[0].x;
        for (int c = 0; c < 4; ++c) r[k][c] = e[4 * c + k];
    }
    auto at_center = [&](const double* q) { return q[0] * m->center[0] + q[1] * m->center[1] + q[2] * m->center[2] + q[3]; };
    auto slope = [&](const double* q) { return std::sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2]); };
    // A linear function of p over the sphere lies within value(centre) +- slope * radius.
    // Float rounding of the kernels' transform and plane distances is far
    // below the margin (1e-4 of the magnitudes involved).
    auto margin = [&](const double* q) { return 1e-4 * (slope(q) * (std::fabs(m->center[0]) + std::fabs(m->center[1]) +
                                                                     std::fabs(m->center[2]) + m->radius) + std::fabs(q[3])); };
    for (int plane = 0; plane < 6; ++plane) {
        // x + w, w - x, y + w, w - y, z + w, w - z (reference pipeline.cpp:274-285)
        const int axis = plane / 2;
        const double sg = (plane % 2) ? -1.0 : 1.0;
        double q[4];
        for (int c = 0; c < 4; ++c) q[c] = r[3][c] + sg * r[axis][c];
        if (!(at_center(q) - slope(q) * m->radius > margin(q))) return s;
    }
    s.clip = true;
    // The back-face pre-test (k_vis): the eye in object space is where rows
    // x, y and w of the transform vanish, A3 e + a = 0 (A3: those rows'
    // first three columns, a: their last).  Then (x, y, w) of a vertex is
    // A3 (p - e), and a face's screen-space cross product is
    // (W H / 4) det(A3) ((p0 - e) . n) / (w0 w1 w2), n = (p1 - p0) x (p2 - p0).
    {
        const double* R[3] = {r[0], r[1], r[3]};
        auto det3 = [](const double (&m)[3][3]) {
            return m[0][0] * (m[1][1] * m[2][2] - m[1][2] * m[2][1]) - m[0][1] * (m[1][0] * m[2][2] - m[1][2] * m[2][0]) +
                   m[0][2] * (m[1][0] * m[2][1] - m[1][1] * m[2][0]);
        };
        double A[3][3];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) A[i][j] = R[i][j];
        const double det = det3(A);
        const double w_hi = at_center(r[3]) + slope(r[3]) * m->radius;
        if (std::isfinite(det) && det != 0.0 && w_hi > 0.0) {
            for (int j = 0; j < 3; ++j) {   // Cramer's rule for A3 e = -a
                double Aj[3][3];
                for (int i = 0; i < 3; ++i)
                    for (int k = 0; k < 3; ++k) Aj[i][k] = k == j ? -R[i][3] : A[i][k];
                s.eye[j] = det3(Aj) / det;
            }
            s.sign = det > 0.0 ? 1.0 : -1.0;
            // |(p0 - e) . n| above tau: the screen cross product exceeds
            // 0.25 px^2 (w <= w_hi over the sphere), some 40 times the float
            // error of the kernels' screen-space test for faces of a few pixels
            s.tau = w_hi * w_hi * w_hi / ((double)W * (double)H * std::fabs(det));
            if (!std::isfinite(s.tau) || !std::isfinite(s.eye[0] + s.eye[1] + s.eye[2])) s.sign = 0.0;
        }
    }
    // Screen extent of any face: for p, q within max_edge of each other in
    // the sphere, |x_ndc(p) - x_ndc(q)| <= max_edge * sup |grad(a.P / b.P)|
    // over the segment, with grad = (a' (b.P) - b' (a.P)) / (b.P)^2.  The
    // segment lies inside the frustum (every point of the sphere is), where
    // |a.P| <= b.P = w, so |grad| <= (|a'| + |b'|) / w <= (|a'| + |b'|) / w_lo.
    const double w_lo = at_center(r[3]) - slope(r[3]) * m->radius;
    if (!(w_lo > 0)) return s;
    auto extent = [&](const double* a, int size) {
        const double g = (slope(a) + slope(r[3])) / w_lo;
        return 0.5 * size * m->max_edge * g * (1.0 + 1e-4) + 1e-3;
    };
    // the pixel rectangle spans floor(min)..ceil(max): at most extent + 2 pixels
    const double ex = std::ceil(extent(r[0], W)) + 2.0, ey = std::ceil(extent(r[1], H)) + 2.0;
    s.big = ex * ey <= (double)footprint;
    return s;
}

int uniform_shader(const ViewDesc* views, uint32_t n) {
    for (uint32_t i = 1; i < n; ++i)
        if (views[i].u.kind != views[0].u.kind) return -1;
    return n ? views[0].u.kind : -1;
}

ViewDesc make_view(const axr_mesh* m, uint32_t* color, float* depth, uint32_t* ids, int w, int h,
                   const axr_draw_params* p, int tile) {
    ViewDesc v{};
    v.pos = m->pos;
    v.verts = m->verts;
    v.idx = m->idx;
    v.face_colors = m->face_colors;
    v.groups = m->groups;
    v.g0 = m->g0;
    v.materials = m->materials;
    v.ngroups = m->ngroups;
    v.nfaces = m->nfaces_sub;
    mat4 model, vp;
    std::memcpy(&model, p->model, 64);
    std::memcpy(&vp, p->view_proj, 64);
    // reference tiled_pipeline.cpp:148-155: mvp = viewProj * model (the
    // viewport matrix is a uniform there too, but no shader reads it)
    v.mvp = vp * model;
    Uniforms& u = v.u;
    u.model = model;
    u.normal_mat = upper3(transpose(inverse(model)));
    u.camera_pos = make3(p->camera_pos[0], p->camera_pos[1], p->camera_pos[2]);
    u.light_dir = make3(p->light_dir[0], p->light_dir[1], p->light_dir[2]);
    u.light_color = make3(p->light_color[0], p->light_color[1], p->light_color[2]);
    u.kind = p->shader;
    u.num_lights = std::max(0, std::min(p->num_lights, MAX_POINT_LIGHTS));
    for (int l = 0; l < MAX_POINT_LIGHTS; ++l) {
        u.light_pos[l] = make3(p->light_pos[l][0], p->light_pos[l][1], p->light_pos[l][2]);
        u.light_colors[l] = make3(p->light_colors[l][0], p->light_colors[l][1], p->light_colors[l][2]);
    }
    v.color = color;
    v.depth = depth;
    v.ids = ids;
    v.W = w;
    v.H = h;
    v.tiles_x = (w + tile - 1) / tile;
    v.tiles_y = (h + tile - 1) / tile;
    return v;
}

bool shader_ok(const axr_draw_params* p) { return p->shader >= 0 && p->shader < SHADER_COUNT; }

}  // namespace axr

namespace {

void harvest_events(axr_ctx* ctx) {
    while (ctx->ev_pending > 0) {
        int slot = (ctx->ev_head - ctx->ev_pending + axr_ctx::kEventRing) % axr_ctx::kEventRing;
        StageEvents& e = ctx->ev[slot];
        --ctx->ev_pending;
        if (!e.used) continue;   // the frame failed before its first kernel
        const int first = __builtin_ctz(e.used), last = 31 - __builtin_clz(e.used);
        if (hipEventSynchronize(e.stop[last]) != hipSuccess) continue;
        float total = 0;
        (void)hipEventElapsedTime(&total, e.start[first], e.stop[last]);
        ctx->stats.frames += 1;
        ctx->stats.frame_ms_total += total;
        for (int k = 0; k < kStages; ++k) {
            if (!((e.used >> k) & 1u)) continue;
            float ms = 0;
            (void)hipEventElapsedTime(&ms, e.start[k], e.stop[k]);
            ctx->stats.stage_ms_total[k] += ms;
        }
    }
}

// Kernel timing events for one frame (null when profiling is off).
StageEvents* next_events(axr_ctx* ctx) {
    if (!ctx->profiling) return nullptr;
    if (ctx->ev_pending == axr_ctx::kEventRing) harvest_events(ctx);
    StageEvents* ev = &ctx->ev[ctx->ev_head];
    ev->used = 0;
    ctx->ev_head = (ctx->ev_head + 1) % axr_ctx::kEventRing;
    ++ctx->ev_pending;
    return ev;
}

int ctx_launch(axr_ctx* ctx, const KArgs& ka) { return arena_launch(ctx->err, ctx->main, ka, next_events(ctx)); }

}  // namespace

extern "C" {

const char* axr_version(void) { return "axiomr-mi355x 0.1 (gfx950)"; }

int axr_check_view(uint64_t submitted_faces, int width, int height) {
    if (submitted_faces >= MAX_VIEW_FACES) return AXR_ERR_INVALID;
    if (width <= 0 || height <= 0 || width > MAX_FB_EDGE || height > MAX_FB_EDGE) return AXR_ERR_INVALID;
    return AXR_OK;
}

int axr_device_count(int* count) {
    if (!count) return AXR_ERR_INVALID;
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    *count = (e == hipSuccess) ? n : 0;
    return AXR_OK;
}

int axr_ctx_create(int device, uint64_t hbm_budget_bytes, axr_ctx** out) {
    if (!out) return AXR_ERR_INVALID;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return AXR_ERR_NO_DEVICE;
    if (device < 0 || device >= n) return AXR_ERR_INVALID;
    axr_ctx* ctx = new (std::nothrow) axr_ctx();
    if (!ctx) return AXR_ERR_OOM;
    ctx->device = device;
    ctx->budget = hbm_budget_bytes;
    ctx->cfg.tile_size = 16;
    ctx->cfg.raster_threads = 64;
    ctx->cfg.bin_capacity = 128;
    ctx->cfg.small_footprint = 12;
    if (hipSetDevice(device) != hipSuccess) { delete ctx; return AXR_ERR_HIP; }
    for (auto* evs : {ctx->join_ev, ctx->sync_ev})
        for (int k = 0; k < axr_ctx::kArenas; ++k)
            if (hipEventCreateWithFlags(&evs[k], hipEventDisableTiming) != hipSuccess) {
                axr_ctx_destroy(ctx);
                return AXR_ERR_HIP;
            }
    if (hipHostMalloc((void**)&ctx->noalloc_counters, CNT_COUNT * 4, hipHostMallocDefault) != hipSuccess) {
        axr_ctx_destroy(ctx);
        return AXR_ERR_OOM;
    }
    int r = arena_create(ctx, ctx->err, ctx->main);
    if (r != AXR_OK) { axr_ctx_destroy(ctx); return r; }
    *out = ctx;
    return AXR_OK;
}

int axr_ctx_destroy(axr_ctx* ctx) {
    if (!ctx) return AXR_ERR_INVALID;
    (void)hipSetDevice(ctx->device);
    // every stream first: a stream may still wait on another arena's signal,
    // which returns to the shared free list when its arena is destroyed
    if (ctx->main.stream) (void)ctx_sync_all(ctx);
    for (auto& a : ctx->alt) arena_destroy(ctx, a);
    arena_destroy(ctx, ctx->main);
    for (auto* evs : {ctx->join_ev, ctx->sync_ev})
        for (int k = 0; k < axr_ctx::kArenas; ++k)
            if (evs[k]) (void)hipEventDestroy(evs[k]);
    if (ctx->noalloc_counters) (void)hipHostFree(ctx->noalloc_counters);
    if (ctx->ev) {
        for (int i = 0; i < axr_ctx::kEventRing; ++i)
            for (int k = 0; k < kStages; ++k) {
                if (ctx->ev[i].start[k]) (void)hipEventDestroy(ctx->ev[i].start[k]);
                if (ctx->ev[i].stop[k]) (void)hipEventDestroy(ctx->ev[i].stop[k]);
            }
        delete[] ctx->ev;
    }
    delete ctx;
    return AXR_OK;
}

const char* axr_last_error(const axr_ctx* ctx) { return ctx ? ctx->err.msg.c_str() : "null context"; }

int axr_ctx_set_config(axr_ctx* ctx, const axr_config* cfg) {
    if (!ctx || !cfg) return AXR_ERR_INVALID;
    if (cfg->tile_size != 8 && cfg->tile_size != 16 && cfg->tile_size != 32)
        return fail(ctx->err, AXR_ERR_INVALID, "tile_size must be 8, 16 or 32 (got %d)", cfg->tile_size);
    if (cfg->raster_threads != 64 && cfg->raster_threads != 128 && cfg->raster_threads != 256)
        return fail(ctx->err, AXR_ERR_INVALID, "raster_threads must be 64, 128 or 256 (got %d)", cfg->raster_threads);
    if (cfg->bin_capacity < 64 || cfg->bin_capacity > 1024)
        return fail(ctx->err, AXR_ERR_INVALID, "bin_capacity must be in [64, 1024] (got %d)", cfg->bin_capacity);
    if (cfg->small_footprint < 1 || cfg->small_footprint > 1024)
        return fail(ctx->err, AXR_ERR_INVALID, "small_footprint must be in [1, 1024] (got %d)", cfg->small_footprint);
    if (cfg->raster_mode != AXR_RASTER_BINNED && cfg->raster_mode != AXR_RASTER_VISIBILITY)
        return fail(ctx->err, AXR_ERR_INVALID, "raster_mode must be AXR_RASTER_BINNED or AXR_RASTER_VISIBILITY (got %d)",
                    cfg->raster_mode);
    ctx->cfg = *cfg;
    return AXR_OK;
}

int axr_ctx_get_config(const axr_ctx* ctx, axr_config* cfg) {
    if (!ctx || !cfg) return AXR_ERR_INVALID;
    *cfg = ctx->cfg;
    return AXR_OK;
}

int axr_ctx_set_profiling(axr_ctx* ctx, int enable) {
    if (!ctx) return AXR_ERR_INVALID;
    (void)hipSetDevice(ctx->device);
    if (enable && !ctx->ev) {
        ctx->ev = new (std::nothrow) StageEvents[axr_ctx::kEventRing];
        if (!ctx->ev) return fail(ctx->err, AXR_ERR_OOM, "profiling events");
        for (int i = 0; i < axr_ctx::kEventRing; ++i)
            for (int k = 0; k < kStages; ++k)
                if (hipEventCreate(&ctx->ev[i].start[k]) != hipSuccess ||
                    hipEventCreate(&ctx->ev[i].stop[k]) != hipSuccess)
                    return fail(ctx->err, AXR_ERR_HIP, "hipEventCreate");
    }
    if (!enable) harvest_events(ctx);
    ctx->profiling = enable != 0;
    ctx->stats = axr_kernel_stats{};
    return AXR_OK;
}

int axr_ctx_get_stats(axr_ctx* ctx, axr_kernel_stats* out) {
    if (!ctx || !out) return AXR_ERR_INVALID;
    (void)hipSetDevice(ctx->device);
    harvest_events(ctx);
    *out = ctx->stats;
    return AXR_OK;
}

int axr_ctx_memory(const axr_ctx* ctx, uint64_t* in_use, uint64_t* arena) {
    if (!ctx) return AXR_ERR_INVALID;
    auto* c = const_cast<axr_ctx*>(ctx);
    std::lock_guard<std::mutex> lk(c->mu);
    if (in_use) *in_use = ctx->in_use;
    if (arena) *arena = ctx->arena_bytes;
    return AXR_OK;
}

// After a host wait and arena_counters_after_wait: an arena whose frames
// overflowed is grown to what the kernels reported (the current capacity is
// the floor) and reset.
static int arena_settle_overflow(axr_ctx* ctx, FrameArena& fa) {
    if (!fa.host_counters[CNT_ERROR]) return AXR_OK;
    FrameSize fs;
    fs.faces = fa.rec.bytes / sizeof(TriRec);
    fs.tiles = fa.tile_count.bytes / 4 - 1;
    fs.pixels = fa.vkeys.bytes / 8;
    fs.wave_cap = fa.a.wave_cap;
    return arena_check_overflow(ctx, ctx->err, fa, fs);
}

int axr_sync(axr_ctx* ctx) {
    if (!ctx) return AXR_ERR_INVALID;
    (void)hipSetDevice(ctx->device);
    // an event behind each arena's queued work, all recorded before the
    // first wait: the streams' completions then overlap (a stream wait per
    // arena in turn costs ~10 us each); the counters come back only from an
    // arena whose kernels flagged an error
    int live[axr_ctx::kArenas];
    int n = 0;
    for (int k = 0; k < axr_ctx::kArenas; ++k)
        if (arena_no(ctx, k).stream) live[n++] = k;
    hipError_t e = hipSuccess;
    for (int i = 0; i < n && e == hipSuccess; ++i) e = hipEventRecord(ctx->sync_ev[live[i]], arena_no(ctx, live[i]).stream);
    for (int i = 0; i < n && e == hipSuccess; ++i) e = hipEventSynchronize(ctx->sync_ev[live[i]]);
    if (e != hipSuccess) return hip_fail(ctx->err, e, "hipEventSynchronize");
    int r = AXR_OK;
    for (int i = 0; i < n; ++i) {
        FrameArena& fa = arena_no(ctx, live[i]);
        int r2 = arena_counters_after_wait(ctx->err, fa);
        if (!r2) r2 = arena_settle_overflow(ctx, fa);
        if (!r) r = r2;
    }
    const int r3 = noalloc_settle(ctx);
    return r ? r : r3;
}

// ---- resources -------------------------------------------------------------

int axr_texture_create(axr_ctx* ctx, const uint8_t* rgba, int w, int h, int filter, axr_texture** out) {
    if (!ctx || !out) return AXR_ERR_INVALID;
    *out = nullptr;
    if (!rgba || w <= 0 || h <= 0 || (uint64_t)w * (uint64_t)h > (1ull << 30))
        return fail(ctx->err, AXR_ERR_INVALID, "texture: bad size %dx%d", w, h);
    (void)hipSetDevice(ctx->device);
    axr_texture* t = new (std::nothrow) axr_texture();
    if (!t) return AXR_ERR_OOM;
    t->ctx = ctx;
    t->bytes = (size_t)w * h * 4;
    void* p;
    int r = dev_alloc(ctx, ctx->err, &p, t->bytes);
    if (r) { delete t; return r; }
    if ((r = upload(ctx->err, ctx->main.stream, p, rgba, t->bytes))) { dev_free(ctx, p, t->bytes); delete t; return r; }
    t->dev = DevTexture{(const uint32_t*)p, w, h, filter ? 1 : 0, 0};
    *out = t;
    return AXR_OK;
}

int axr_texture_destroy(axr_texture* t) {
    if (!t) return AXR_ERR_INVALID;
    (void)hipSetDevice(t->ctx->device);
    (void)ctx_sync_all(t->ctx);
    dev_free(t->ctx, (void*)t->dev.texels, t->bytes);
    delete t;
    return AXR_OK;
}

int axr_mesh_create(axr_ctx* ctx, const float* verts56, uint64_t nverts, const uint32_t* idx, uint64_t nfaces,
                    const axr_group* groups, uint32_t ngroups, const axr_material* materials, uint32_t nmaterials,
                    const uint32_t* face_colors, axr_mesh** out) {
    if (!ctx || !out) return AXR_ERR_INVALID;
    *out = nullptr;
    ErrSink& err = ctx->err;
    if (!verts56 || !idx || nverts == 0 || nfaces == 0 || nverts > 0xFFFFFFFFull || nfaces > 0x0FFFFFFFull)
        return fail(err, AXR_ERR_INVALID, "mesh: bad sizes (%llu vertices, %llu faces)", (unsigned long long)nverts,
                    (unsigned long long)nfaces);
    if (!groups || ngroups == 0 || !materials || nmaterials == 0 || nmaterials > REC_MATERIAL_MASK)
        return fail(err, AXR_ERR_INVALID, "mesh: needs at least one group and one material");
    for (uint64_t i = 0; i < 3 * nfaces; ++i)
        if (idx[i] >= nverts)
            return fail(err, AXR_ERR_INVALID, "mesh: index %u at %llu out of range", idx[i], (unsigned long long)i);
    std::vector<DevGroup> dg(ngroups);
    uint64_t rank = 0;
    for (uint32_t g = 0; g < ngroups; ++g) {
        const axr_group& s = groups[g];
        if ((uint64_t)s.first_face + s.face_count > nfaces)
            return fail(err, AXR_ERR_INVALID, "mesh: group %u covers faces past the end", g);
        // reference Mesh::getMaterial uses .at() and throws (include/mesh.hpp:80-82)
        if (s.material >= nmaterials)
            return fail(err, AXR_ERR_INVALID, "mesh: group %u names material %u of %u", g, s.material, nmaterials);
        dg[g] = DevGroup{s.first_face, s.face_count, (uint32_t)rank, s.material};
        rank += s.face_count;
    }
    if (rank >= MAX_VIEW_FACES)
        return fail(err, AXR_ERR_INVALID, "mesh: %llu faces in submission order (the limit is 2^29 - 1)",
                    (unsigned long long)rank);
    std::vector<DevMaterial> dm(nmaterials);
    for (uint32_t i = 0; i < nmaterials; ++i) {
        const axr_material& s = materials[i];
        auto tex = [](const axr_texture* t) { return t ? t->dev : DevTexture{nullptr, 0, 0, 0, 0}; };
        dm[i] = DevMaterial{tex(s.diffuse), tex(s.metallic), tex(s.bump), tex(s.roughness), tex(s.ao),
                            s.specular_exponent, {0, 0, 0}};
    }
    (void)hipSetDevice(ctx->device);
    axr_mesh* m = new (std::nothrow) axr_mesh();
    if (!m) return AXR_ERR_OOM;
    m->ctx = ctx;
    m->nverts = nverts;
    m->nfaces = nfaces;
    m->ngroups = ngroups;
    m->nmaterials = nmaterials;
    m->nfaces_sub = (uint32_t)rank;
    m->g0 = dg[0];
    std::vector<float> pos(3 * nverts);
    for (uint64_t i = 0; i < nverts; ++i)
        for (int k = 0; k < 3; ++k) pos[3 * i + k] = verts56[14 * i + k];
    const size_t sizes[] = {pos.size() * 4, nverts * 56, nfaces * 12, face_colors ? nfaces * 4 : 0,
                            ngroups * sizeof(DevGroup), nmaterials * sizeof(DevMaterial)};
    void** ptrs[] = {(void**)&m->pos, (void**)&m->verts, (void**)&m->idx, (void**)&m->face_colors,
                     (void**)&m->groups, (void**)&m->materials};
    const void* srcs[] = {pos.data(), verts56, idx, face_colors, dg.data(), dm.data()};
    int r = AXR_OK;
    for (int i = 0; i < 6 && r == AXR_OK; ++i) {
        r = dev_alloc(ctx, err, ptrs[i], sizes[i]);
        if (r == AXR_OK) r = upload(err, ctx->main.stream, *ptrs[i], srcs[i], sizes[i]);
    }
    if (r != AXR_OK) {
        for (int i = 0; i < 6; ++i) dev_free(ctx, *ptrs[i], *ptrs[i] ? sizes[i] : 0);
        delete m;
        return r;
    }
    for (size_t s : sizes) m->bytes += s;
    // bounds for pass_skips: centre of the axis-aligned box, radius to the
    // farthest vertex, longest edge
    double lo[3] = {verts56[0], verts56[1], verts56[2]}, hi[3] = {lo[0], lo[1], lo[2]};
    for (uint64_t i = 0; i < nverts; ++i)
        for (int k = 0; k < 3; ++k) {
            lo[k] = std::min(lo[k], (double)verts56[14 * i + k]);
            hi[k] = std::max(hi[k], (double)verts56[14 * i + k]);
        }
    for (int k = 0; k < 3; ++k) m->center[k] = 0.5 * (lo[k] + hi[k]);
    double r2 = 0, e2 = 0;
    for (uint64_t i = 0; i < nverts; ++i) {
        double d = 0;
        for (int k = 0; k < 3; ++k) d += (verts56[14 * i + k] - m->center[k]) * (verts56[14 * i + k] - m->center[k]);
        r2 = std::max(r2, d);
    }
    for (uint64_t f = 0; f < nfaces; ++f)
        for (int k = 0; k < 3; ++k) {
            const float* a = verts56 + 14 * (uint64_t)idx[3 * f + k];
            const float* b = verts56 + 14 * (uint64_t)idx[3 * f + (k + 1) % 3];
            const double dx = (double)a[0] - b[0], dy = (double)a[1] - b[1], dz = (double)a[2] - b[2];
            e2 = std::max(e2, dx * dx + dy * dy + dz * dz);
        }
    m->radius = std::sqrt(r2);
    m->max_edge = std::sqrt(e2);
    if (!std::isfinite(m->radius) || !std::isfinite(m->max_edge)) m->radius = m->max_edge = INFINITY;
    *out = m;
    return AXR_OK;
}

int axr_mesh_destroy(axr_mesh* m) {
    if (!m) return AXR_ERR_INVALID;
    if (m->batch_refs)
        return fail(m->ctx->err, AXR_ERR_INVALID, "mesh is drawn by %u live batch view(s): destroy the batch first",
                    m->batch_refs);
    (void)hipSetDevice(m->ctx->device);
    (void)ctx_sync_all(m->ctx);
    dev_free(m->ctx, m->pos, m->nverts * 12);
    dev_free(m->ctx, m->verts, m->nverts * 56);
    dev_free(m->ctx, m->idx, m->nfaces * 12);
    dev_free(m->ctx, m->face_colors, m->face_colors ? m->nfaces * 4 : 0);
    dev_free(m->ctx, m->groups, m->ngroups * sizeof(DevGroup));
    dev_free(m->ctx, m->materials, m->nmaterials * sizeof(DevMaterial));
    delete m;
    return AXR_OK;
}

int axr_fb_create(axr_ctx* ctx, int w, int h, int with_ids, axr_fb** out) {
    if (!ctx || !out) return AXR_ERR_INVALID;
    *out = nullptr;
    if (w <= 0 || h <= 0 || w > MAX_FB_EDGE || h > MAX_FB_EDGE)
        return fail(ctx->err, AXR_ERR_INVALID, "framebuffer: bad size %dx%d", w, h);
    (void)hipSetDevice(ctx->device);
    axr_fb* fb = new (std::nothrow) axr_fb();
    if (!fb) return AXR_ERR_OOM;
    fb->ctx = ctx;
    fb->w = w;
    fb->h = h;
    const size_t n = (size_t)w * h;
    int r;
    if ((r = dev_alloc(ctx, ctx->err, (void**)&fb->color, n * 4)) ||
        (r = dev_alloc(ctx, ctx->err, (void**)&fb->depth, n * 4)) ||
        (with_ids && (r = dev_alloc(ctx, ctx->err, (void**)&fb->ids, n * 4)))) {
        dev_free(ctx, fb->color, fb->color ? n * 4 : 0);
        dev_free(ctx, fb->depth, fb->depth ? n * 4 : 0);
        delete fb;
        return r;
    }
    fb->bytes = n * (with_ids ? 12 : 8);
    // reference Framebuffer ctor: colour 0, depth +inf (src/framebuffer.cpp:6-14)
    hipError_t e = launch_clear(fb->color, fb->depth, fb->ids, n, 0u, __builtin_huge_valf(), ctx->main.stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->main.stream);
    if (e != hipSuccess) { axr_fb_destroy(fb); return hip_fail(ctx->err, e, "framebuffer init"); }
    *out = fb;
    return AXR_OK;
}

int axr_fb_destroy(axr_fb* fb) {
    if (!fb) return AXR_ERR_INVALID;
    if (fb->batch_refs)
        return fail(fb->ctx->err, AXR_ERR_INVALID, "framebuffer is a view of a live batch: destroy the batch first");
    (void)hipSetDevice(fb->ctx->device);
    (void)ctx_sync_all(fb->ctx);
    const size_t n = (size_t)fb->w * fb->h;
    dev_free(fb->ctx, fb->color, n * 4);
    dev_free(fb->ctx, fb->depth, n * 4);
    dev_free(fb->ctx, fb->ids, fb->ids ? n * 4 : 0);
    delete fb;
    return AXR_OK;
}

int axr_fb_clear(axr_fb* fb, uint32_t bgra, float depth) {
    if (!fb) return AXR_ERR_INVALID;
    // recorded, not written: the next draw in visibility mode writes it as
    // part of its last pass; any other use of the planes writes it first
    fb->clear_pending = true;
    fb->clear_bgra = bgra;
    fb->clear_depth = depth;
    return AXR_OK;
}

int axr_fb_read(axr_fb* fb, uint8_t* bgra, float* depth, uint32_t* ids) {
    if (!fb) return AXR_ERR_INVALID;
    axr_ctx* ctx = fb->ctx;
    (void)hipSetDevice(ctx->device);
    const size_t n = (size_t)fb->w * fb->h;
    int r = fb_settle(fb);
    if (r || (r = ctx_sync_all(ctx))) return r;
    if (ids && !fb->ids) return fail(ctx->err, AXR_ERR_INVALID, "framebuffer has no id plane");
    // copies on the context's own stream: a copy on the null stream would
    // give the process one more hardware queue to share with the frame
    // arenas' streams, which then slows every later overlapped draw
    const hipStream_t st = ctx->main.stream;
    hipError_t e = hipSuccess;
    if (e == hipSuccess && bgra) e = hipMemcpyAsync(bgra, fb->color, n * 4, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess && depth) e = hipMemcpyAsync(depth, fb->depth, n * 4, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess && ids) e = hipMemcpyAsync(ids, fb->ids, n * 4, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    return e == hipSuccess ? AXR_OK : hip_fail(ctx->err, e, "framebuffer read");
}

int axr_texture_load(axr_ctx* ctx, const char* path, int filter, axr_texture** out) {
    if (!ctx || !out) return AXR_ERR_INVALID;
    *out = nullptr;
    int w = 0, h = 0;
    uint8_t* rgba = nullptr;
    int r = axr_image_load(path, &w, &h, &rgba);
    if (r) return fail(ctx->err, r, "texture: %s", axr_image_last_error());
    r = axr_texture_create(ctx, rgba, w, h, filter, out);
    axr_image_free(rgba);
    return r;
}

int axr_fb_write_png(axr_fb* fb, const char* path) {
    if (!fb || !path) return AXR_ERR_INVALID;
    const size_t n = (size_t)fb->w * fb->h;
    std::vector<uint32_t> bgra(n);
    int r = axr_fb_read(fb, reinterpret_cast<uint8_t*>(bgra.data()), nullptr, nullptr);
    if (r) return r;
    std::vector<uint8_t> rgba(4 * n);
    for (size_t i = 0; i < n; ++i) {
        const uint32_t p = bgra[i];
        rgba[4 * i] = (uint8_t)(p >> 16);
        rgba[4 * i + 1] = (uint8_t)(p >> 8);
        rgba[4 * i + 2] = (uint8_t)p;
        rgba[4 * i + 3] = (uint8_t)(p >> 24);
    }
    // framebuffer row 0 is the bottom row; the file's first row is the top
    r = axr_image_write_png(path, fb->w, fb->h, rgba.data(), 1);
    return r ? fail(fb->ctx->err, r, "png: %s", axr_image_last_error()) : AXR_OK;
}

// ---- drawMesh ----------------------------------------------------------------

// An asynchronous visibility-mode draw on the arena whose turn it is: its
// setup (and, for a two-pass draw, shading) passes go on that arena's
// stream at once; its framebuffer pass waits for everything queued so far
// on the stream of the latest framebuffer work (among it the previous
// draw's framebuffer pass), so it overlaps earlier frames and still writes
// framebuffers in call order.
static int draw_pipelined(axr_ctx* ctx, axr_fb* fb, const KArgs& ka, const FrameSize& fs, StageEvents* ev) {
    const int x = ctx->next_arena;
    FrameArena& fa = arena_no(ctx, x);
    int r;
    if (x && !fa.stream && (r = arena_create(ctx, ctx->err, fa))) return r;
    if ((r = arena_ensure(ctx, ctx->err, fa, fs, ctx->need_floor))) return r;
    FbWait wait;
    if (ctx->last_arena != x) {
        FrameArena& last = arena_no(ctx, ctx->last_arena);
        if (ctx->last_value && last.sig) {
            wait.sig = last.sig;
            wait.value = ctx->last_value;
        } else {
            wait.ev = ctx->join_ev[ctx->last_arena];
            hipError_t e = hipEventRecord(wait.ev, last.stream);
            if (e != hipSuccess) return hip_fail(ctx->err, e, "stream order");
        }
    }
    if (x) ctx->alt_unjoined[x - 1] = true;   // before the launch: a failed one may have queued work
    if ((r = arena_launch(ctx->err, fa, ka, ev, wait))) return r;
    fb->clear_pending = false;   // written by this launch, even by a frame that overflows
    // the end of this draw's framebuffer pass, for the next draw to wait on
    // (with last_value 0 the next draw waits with an event on this stream)
    ctx->last_arena = x;
    ctx->next_arena = (x + 1) % axr_ctx::kArenas;
    ctx->last_value = 0;
    if (fa.sig) {
        const uint64_t v = ++g_sig_value;
        hipError_t e = hipStreamWriteValue64(fa.stream, fa.sig, v, 0);
        if (e != hipSuccess) return hip_fail(ctx->err, e, "stream order");
        ctx->last_value = v;
    }
    return AXR_OK;
}

int axr_draw(axr_ctx* ctx, axr_fb* fb, const axr_mesh* mesh, const axr_draw_params* p) {
    if (!ctx) return AXR_ERR_INVALID;
    // reference drawMesh returns silently without a shader/camera/framebuffer
    // (tiled_pipeline.cpp:146); over the C ABI a missing argument is an error.
    if (!fb || !mesh || !p) return fail(ctx->err, AXR_ERR_INVALID, "axr_draw: null framebuffer, mesh or params");
    if (fb->ctx != ctx || mesh->ctx != ctx)
        return fail(ctx->err, AXR_ERR_INVALID, "axr_draw: handle from another context");
    if (!shader_ok(p)) return fail(ctx->err, AXR_ERR_INVALID, "unknown shader %d", p->shader);
    if (axr_check_view(mesh->nfaces_sub, fb->w, fb->h))
        return fail(ctx->err, AXR_ERR_INVALID, "axr_draw: view too large (%u faces, %dx%d)", mesh->nfaces_sub, fb->w,
                    fb->h);
    (void)hipSetDevice(ctx->device);
    const int tile = ctx->cfg.tile_size;
    KArgs ka{};
    ka.nviews = 1;
    ka.one = make_view(mesh, fb->color, fb->depth, fb->ids, fb->w, fb->h, p, tile);
    ka.total_faces = mesh->nfaces_sub;
    ka.total_tiles = (uint32_t)(ka.one.tiles_x * ka.one.tiles_y);
    ka.tile_size = tile;
    ka.raster_threads = ctx->cfg.raster_threads;
    ka.small_footprint = ctx->cfg.small_footprint;
    ka.mode = ctx->cfg.raster_mode;
    ka.total_pixels = (uint32_t)((uint64_t)fb->w * (uint64_t)fb->h);
    ka.shader = p->shader;
    if (!ctx->no_pass_skips) {
        const PassSkips sk = pass_skips(mesh, ka.one.mvp, fb->w, fb->h, ka.small_footprint);
        ka.skip_clip = sk.clip;
        ka.skip_big = sk.big && ka.mode == RASTER_VISIBILITY;
        if (sk.clip && sk.sign != 0.0) {
            for (int j = 0; j < 3; ++j) ka.cull_eye[j] = (float)sk.eye[j];
            ka.cull_sign = (float)sk.sign;
            ka.cull_tau = (float)sk.tau;
        }
        ctx->stats.passes_skipped += (uint64_t)ka.skip_clip + (uint64_t)ka.skip_big;
    }
    if (fb->clear_pending && ka.mode == RASTER_VISIBILITY) {
        // the resolve pass writes the clear (k_resolve, ka.clear)
        ka.clear = 1;
        ka.clear_bgra = fb->clear_bgra;
        ka.clear_depth = fb->clear_depth;
    }
    const bool async = (p->flags & AXR_DRAW_ASYNC) != 0;
    // an asynchronous draw that absorbs a clear shades apart from its
    // framebuffer writes (also when profiled, so the stage times are those
    // of the same kernels); with an id plane the winners' colours wait in
    // the arena's colour slots
    ka.two_pass = async && ka.clear;
    FrameSize fs = frame_size(ka, (uint32_t)ctx->cfg.bin_capacity);
    if (ka.two_pass && fb->ids) fs.color_slots = fs.pixels;
    if (async && ka.mode == RASTER_VISIBILITY && overlap_enabled()) return draw_pipelined(ctx, fb, ka, fs, next_events(ctx));
    int r;
    if (fb->clear_pending && !ka.clear) {
        // binned mode: the clear is written now, so the winners are tested
        // against its depth without loading the plane
        if ((r = fb_settle(fb))) return r;
        ka.depth_is_clear = 1;
        ka.clear_depth = fb->clear_depth;
    }
    if ((r = ctx_join(ctx)) || (r = arena_ensure(ctx, ctx->err, ctx->main, fs, ctx->need_floor))) return r;
    for (int attempt = 0; attempt < 8; ++attempt) {
        if ((r = ctx_launch(ctx, ka))) return r;
        fb->clear_pending = false;   // written by this launch, even by a frame that overflows
        if (async) return AXR_OK;
        const hipError_t e = hipStreamSynchronize(ctx->main.stream);
        if (e != hipSuccess) return hip_fail(ctx->err, e, "hipStreamSynchronize");
        if ((r = arena_counters_after_wait(ctx->err, ctx->main))) return r;
        if (!ctx->main.host_counters[CNT_ERROR]) return AXR_OK;
        r = arena_check_overflow(ctx, ctx->err, ctx->main, fs);
        if (r != AXR_ERR_OVERFLOW) return r;
        // the raster stage was skipped, so the framebuffer is untouched: redraw
    }
    return fail(ctx->err, AXR_ERR_OVERFLOW, "axr_draw: frame arena kept overflowing");
}

// ---- batched form --------------------------------------------------------------

// The uniforms a vertex stage reads are equal: the model matrix (normal-mapped
// and Gouraud shaders), the normal matrix (flat shader), the light direction
// (Gouraud).
static bool same_vertex_uniforms(const Uniforms& a, const Uniforms& b) {
    return std::memcmp(&a.model, &b.model, sizeof(a.model)) == 0 &&
           std::memcmp(&a.normal_mat, &b.normal_mat, sizeof(a.normal_mat)) == 0 &&
           std::memcmp(&a.light_dir, &b.light_dir, sizeof(a.light_dir)) == 0;
}

// Groups the batch's views into vertex-stage jobs (views of one mesh with
// one shader kind and equal vertex-stage uniforms share one), allocates each
// job's output and points the views at it.  The batch's launch runs k_vary
// over all jobs before the resolve reads them.
static int batch_plan_vary(axr_batch* b, const axr_mesh* const* meshes) {
    uint64_t total = 0;
    for (uint32_t i = 0; i < b->nviews; ++i) {
        ViewDesc& v = b->views[i];
        const uint32_t vb = vary_bytes(v.u.kind);
        if (!vb) continue;
        size_t j = 0;
        for (; j < b->vary_jobs.size(); ++j) {
            const VaryJob& q = b->vary_jobs[j];
            if (q.verts == v.verts && q.kind == v.u.kind && same_vertex_uniforms(b->views[q.view].u, v.u)) break;
        }
        if (j == b->vary_jobs.size()) {
            const uint64_t nv = meshes[i]->nverts;
            if (total + nv > 0x7FFFFFFFull) continue;   // the rest run their vertex stage per pixel
            void* out = nullptr;
            const size_t bytes = (size_t)nv * vb;
            int r = dev_alloc(b->ctx, b->ctx->err, &out, bytes);
            if (r) return r;
            b->vary_jobs.push_back(VaryJob{v.verts, out, (uint32_t)total, (uint32_t)nv, i, v.u.kind});
            b->vary_bytes.push_back(bytes);
            total += nv;
        }
        v.vary = b->vary_jobs[j].out;
    }
    return AXR_OK;
}

int axr_batch_create(axr_ctx* ctx, uint32_t nviews, axr_fb* const* fbs, const axr_mesh* const* meshes,
                     const axr_draw_params* params, axr_batch** out) {
    if (!ctx || !out) return AXR_ERR_INVALID;
    *out = nullptr;
    if (nviews == 0 || !fbs || !meshes || !params)
        return fail(ctx->err, AXR_ERR_INVALID, "batch: empty or null arrays");
    (void)hipSetDevice(ctx->device);
    axr_batch* b = new (std::nothrow) axr_batch();
    if (!b) return AXR_ERR_OOM;
    b->ctx = ctx;
    b->nviews = nviews;
    b->views.resize(nviews);
    b->fbs.assign(fbs, fbs + nviews);
    std::vector<uint32_t> fbase(nviews + 1), tbase(nviews + 1), pbase(nviews + 1);
    uint64_t faces = 0, tiles = 0, pixels = 0;
    const int tile = ctx->cfg.tile_size;
    bool all_inside = true;
    // every view writes its framebuffer in the same launch, with no order
    // between views: two views of one framebuffer would race on its pixels
    // (the reference's render loop draws meshes into one framebuffer one
    // drawMesh after the other, renderer.cpp:100-110; here that is one view
    // per framebuffer, or one batch per mesh)
    std::vector<const axr_fb*> seen(fbs, fbs + nviews);
    std::sort(seen.begin(), seen.end());
    if (std::adjacent_find(seen.begin(), seen.end()) != seen.end()) {
        delete b;
        return fail(ctx->err, AXR_ERR_INVALID, "batch: a framebuffer appears in more than one view");
    }
    for (uint32_t i = 0; i < nviews; ++i) {
        if (!fbs[i] || !meshes[i] || fbs[i]->ctx != ctx || meshes[i]->ctx != ctx || !shader_ok(&params[i])) {
            delete b;
            return fail(ctx->err, AXR_ERR_INVALID, "batch: view %u has a bad handle or shader", i);
        }
        if (axr_check_view(meshes[i]->nfaces_sub, fbs[i]->w, fbs[i]->h)) {
            delete b;
            return fail(ctx->err, AXR_ERR_INVALID, "batch: view %u too large (%u faces)", i, meshes[i]->nfaces_sub);
        }
        const axr_fb* fb = fbs[i];
        ViewDesc v = make_view(meshes[i], fb->color, fb->depth, fb->ids, fb->w, fb->h, &params[i], tile);
        v.inside = !ctx->no_pass_skips && pass_skips(meshes[i], v.mvp, fb->w, fb->h, ctx->cfg.small_footprint).clip;
        all_inside = all_inside && v.inside;
        v.face_base = (uint32_t)faces;
        v.tile_base = (uint32_t)tiles;
        v.pix_base = (uint32_t)pixels;
        fbase[i] = (uint32_t)faces;
        tbase[i] = (uint32_t)tiles;
        pbase[i] = (uint32_t)pixels;
        faces += v.nfaces;
        tiles += (uint64_t)v.tiles_x * v.tiles_y;
        pixels += (uint64_t)v.W * v.H;
        if (faces > MAX_BATCH_FACES || tiles > 0x7FFFFFFFull || pixels > 0xFFFFFFFFull) {
            delete b;
            return fail(ctx->err, AXR_ERR_INVALID, "batch too large");
        }
        b->views[i] = v;
    }
    fbase[nviews] = (uint32_t)faces;
    tbase[nviews] = (uint32_t)tiles;
    pbase[nviews] = (uint32_t)pixels;
    b->total_pixels = (uint32_t)pixels;
    int r;
    if ((r = batch_plan_vary(b, meshes))) {
        axr_batch_destroy(b);
        return r;
    }
    std::vector<uint32_t> tview(tiles);
    for (uint32_t i = 0; i < nviews; ++i)
        for (uint32_t t = tbase[i]; t < tbase[i + 1]; ++t) tview[t] = i;
    const size_t sz_v = nviews * sizeof(ViewDesc), sz_b = (nviews + 1) * 4;
    hipStream_t st = ctx->main.stream;
    ErrSink& err = ctx->err;
    KArgs& ka = b->ka;
    ka.total_tiles = (uint32_t)tiles;
    if ((r = dev_alloc(ctx, err, (void**)&b->d_views, sz_v)) ||
        (r = dev_alloc(ctx, err, (void**)&b->d_face_bases, sz_b)) ||
        (r = dev_alloc(ctx, err, (void**)&b->d_tile_bases, sz_b)) ||
        (r = dev_alloc(ctx, err, (void**)&b->d_pix_bases, sz_b)) ||
        (r = dev_alloc(ctx, err, (void**)&b->d_tile_views, tiles * 4)) ||
        (r = upload(err, st, b->d_views, b->views.data(), sz_v)) ||
        (r = upload(err, st, b->d_face_bases, fbase.data(), sz_b)) ||
        (r = upload(err, st, b->d_tile_bases, tbase.data(), sz_b)) ||
        (r = upload(err, st, b->d_pix_bases, pbase.data(), sz_b)) ||
        (r = upload(err, st, b->d_tile_views, tview.data(), tiles * 4))) {
        axr_batch_destroy(b);
        return r;
    }
    b->bytes = sz_v + 3 * sz_b + tiles * 4;
    if (!b->vary_jobs.empty()) {
        const uint32_t nj = (uint32_t)b->vary_jobs.size();
        std::vector<uint32_t> firsts(nj);
        for (uint32_t j = 0; j < nj; ++j) firsts[j] = b->vary_jobs[j].first;
        if ((r = dev_alloc(ctx, err, (void**)&b->d_vary_jobs, nj * sizeof(VaryJob))) ||
            (r = dev_alloc(ctx, err, (void**)&b->d_vary_firsts, nj * 4)) ||
            (r = upload(err, st, b->d_vary_jobs, b->vary_jobs.data(), nj * sizeof(VaryJob))) ||
            (r = upload(err, st, b->d_vary_firsts, firsts.data(), nj * 4))) {
            axr_batch_destroy(b);
            return r;
        }
        ka.vary_jobs = b->d_vary_jobs;
        ka.vary_firsts = b->d_vary_firsts;
        ka.vary_njobs = nj;
        ka.vary_total = b->vary_jobs.back().first + b->vary_jobs.back().count;
    }
    ka.table = b->d_views;
    ka.face_bases = b->d_face_bases;
    ka.tile_bases = b->d_tile_bases;
    ka.tile_views = b->d_tile_views;
    ka.nviews = nviews;
    ka.total_faces = (uint32_t)faces;
    ka.tile_size = tile;
    ka.raster_threads = ctx->cfg.raster_threads;
    ka.small_footprint = ctx->cfg.small_footprint;
    ka.mode = ctx->cfg.raster_mode;
    ka.total_pixels = (uint32_t)pixels;
    ka.shader = uniform_shader(b->views.data(), nviews);
    ka.one = b->views[0];
    // binned mode: every view inside its frustum, so neither the setup's
    // classification nor the clipper is run
    ka.skip_clip = all_inside && ka.mode != RASTER_VISIBILITY;
    b->fs = frame_size(ka, (uint32_t)ctx->cfg.bin_capacity);
    if ((r = arena_ensure(ctx, err, ctx->main, b->fs))) { axr_batch_destroy(b); return r; }
    // the batch's tables and captured graph hold the views' device pointers:
    // their framebuffers and meshes cannot be destroyed before the batch
    for (uint32_t i = 0; i < nviews; ++i) {
        ++fbs[i]->batch_refs;
        ++const_cast<axr_mesh*>(meshes[i])->batch_refs;
    }
    b->meshes.assign(meshes, meshes + nviews);
    b->holds_refs = true;
    *out = b;
    return AXR_OK;
}

int axr_batch_launch(axr_batch* b) {
    if (!b) return AXR_ERR_INVALID;
    axr_ctx* ctx = b->ctx;
    (void)hipSetDevice(ctx->device);
    FrameArena& fa = ctx->main;
    int r = arena_ensure(ctx, ctx->err, fa, b->fs);
    if (r) return r;
    // a pending clear is overwritten by the batch's own clear, else written first
    for (axr_fb* fb : b->fbs) {
        if (b->clear) fb->clear_pending = false;
        else if ((r = fb_settle(fb))) return r;
    }
    if ((r = ctx_join(ctx))) return r;
    // in visibility mode the resolve pass writes the batch's clear
    KArgs ka = b->ka;
    const bool absorb = b->clear && ka.mode == RASTER_VISIBILITY;
    if (absorb) {
        ka.clear = 1;
        ka.clear_bgra = b->clear_bgra;
        ka.clear_depth = b->clear_depth;
    }
    // k_clear_views writes the clear just before the frame's kernels (writing
    // it from the setup grid instead, beside the face work, took the setup
    // from 0.50 to 0.91 ms on the 256-view batch: slower than the 0.2 ms pass)
    const bool clear_pass = b->clear && !absorb;
    if (clear_pass) {
        ka.depth_is_clear = 1;
        ka.clear_depth = b->clear_depth;
    }
    if (ctx->profiling) {   // events need the eager path
        if (clear_pass) {
            hipError_t e = launch_clear_views(b->d_views, b->d_pix_bases, b->nviews, b->total_pixels, b->clear_bgra,
                                              b->clear_depth, fa.stream);
            if (e != hipSuccess) return hip_fail(ctx->err, e, "batch clear");
        }
        return ctx_launch(ctx, ka);
    }
    if (fa.dirty) {
        hipError_t e = launch_prep(fa.a, (uint32_t)(fa.tile_count.bytes / 4), fa.stream);
        if (e != hipSuccess) return hip_fail(ctx->err, e, "arena reset");
        fa.dirty = false;
    }
    if (!b->exec || b->captured_gen != fa.gen || b->captured_clear != b->clear_version) {
        if (b->exec) { (void)hipGraphExecDestroy(b->exec); b->exec = nullptr; }
        if (b->graph) { (void)hipGraphDestroy(b->graph); b->graph = nullptr; }
        hipError_t e = hipStreamBeginCapture(fa.stream, hipStreamCaptureModeThreadLocal);
        if (e != hipSuccess) return hip_fail(ctx->err, e, "hipStreamBeginCapture");
        hipError_t le = hipSuccess;
        if (clear_pass)
            le = launch_clear_views(b->d_views, b->d_pix_bases, b->nviews, b->total_pixels, b->clear_bgra,
                                    b->clear_depth, fa.stream);
        if (le == hipSuccess) le = launch_frame(ka, fa.a, fa.stream, nullptr);
        e = hipStreamEndCapture(fa.stream, &b->graph);
        if (le != hipSuccess) return hip_fail(ctx->err, le, "batch capture launch");
        if (e != hipSuccess) return hip_fail(ctx->err, e, "hipStreamEndCapture");
        e = hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0);
        if (e != hipSuccess) return hip_fail(ctx->err, e, "hipGraphInstantiate");
        b->captured_gen = fa.gen;
        b->captured_clear = b->clear_version;
    }
    hipError_t e = hipGraphLaunch(b->exec, fa.stream);
    return e == hipSuccess ? AXR_OK : hip_fail(ctx->err, e, "hipGraphLaunch");
}

int axr_batch_set_clear(axr_batch* b, int enable, uint32_t bgra, float depth) {
    if (!b) return AXR_ERR_INVALID;
    b->clear = enable != 0;
    b->clear_bgra = bgra;
    b->clear_depth = depth;
    ++b->clear_version;
    return AXR_OK;
}

int axr_batch_destroy(axr_batch* b) {
    if (!b) return AXR_ERR_INVALID;
    axr_ctx* ctx = b->ctx;
    (void)hipSetDevice(ctx->device);
    (void)ctx_sync_all(ctx);
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    const size_t sz_v = b->nviews * sizeof(ViewDesc), sz_b = (b->nviews + 1) * 4;
    dev_free(ctx, b->d_views, sz_v);
    dev_free(ctx, b->d_face_bases, sz_b);
    dev_free(ctx, b->d_tile_bases, sz_b);
    dev_free(ctx, b->d_pix_bases, sz_b);
    dev_free(ctx, b->d_tile_views, (size_t)b->ka.total_tiles * 4);
    for (size_t j = 0; j < b->vary_jobs.size(); ++j) dev_free(ctx, b->vary_jobs[j].out, b->vary_bytes[j]);
    dev_free(ctx, b->d_vary_jobs, b->vary_jobs.size() * sizeof(VaryJob));
    dev_free(ctx, b->d_vary_firsts, b->vary_jobs.size() * 4);
    if (b->holds_refs) {
        for (axr_fb* fb : b->fbs) --fb->batch_refs;
        for (const axr_mesh* m : b->meshes) --const_cast<axr_mesh*>(m)->batch_refs;
    }
    delete b;
    return AXR_OK;
}

}  // extern "C"
