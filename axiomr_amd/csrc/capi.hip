// capi.hip - host side of the C ABI declared in include/axiomr.h: context,
// device resources, the frame arena, drawMesh, the graph-captured batch.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>
#include "../../include/axiomr.h"
#include "axr_types.hpp"

namespace axr {
hipError_t launch_frame(const KArgs& ka, const Arena& ar, hipStream_t st, hipEvent_t* ev);
hipError_t launch_clear(uint32_t* color, float* depth, uint32_t* ids, size_t n, uint32_t bgra, float d, hipStream_t st);
hipError_t launch_prep(const Arena& ar, uint32_t ntiles, hipStream_t st);
hipError_t launch_clear_views(const ViewDesc* views, const uint32_t* pix_bases, uint32_t nviews, uint32_t total,
                              uint32_t bgra, float d, hipStream_t st);
}

using namespace axr;

namespace {

constexpr int kEventRing = 64;

struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
};

}  // namespace

struct axr_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    uint64_t budget = 0;          // 0 = unlimited
    uint64_t in_use = 0;          // resource bytes (meshes, textures, framebuffers, batches)
    uint64_t arena_bytes = 0;
    std::string err;
    axr_config cfg{};
    // frame arena
    DevBuf counters, rec, clip_faces, clip_rec, clip_verts, tile_count, tile_off, pairs, wave_pairs, bins, active,
        scan_sums;
    bool arena_dirty = true;   // bin counts / counters need a k_prep before the next frame
    Arena arena{};
    uint64_t arena_gen = 0;
    uint32_t* host_counters = nullptr;   // pinned, CNT_COUNT words
    // profiling
    bool profiling = false;
    hipEvent_t ev[kEventRing][6] = {};
    int ev_head = 0, ev_pending = 0;
    axr_kernel_stats stats{};
};

struct axr_texture {
    axr_ctx* ctx;
    DevTexture dev;
    size_t bytes;
};

struct axr_mesh {
    axr_ctx* ctx;
    float* pos = nullptr;
    Vertex56* verts = nullptr;
    uint32_t* idx = nullptr;
    uint32_t* face_colors = nullptr;
    DevGroup* groups = nullptr;
    DevMaterial* materials = nullptr;
    uint32_t ngroups = 0, nmaterials = 0;
    uint64_t nverts = 0, nfaces = 0;
    uint32_t nfaces_sub = 0;   // faces in submission order
    size_t bytes = 0;
};

struct axr_fb {
    axr_ctx* ctx;
    int w, h;
    uint32_t* color = nullptr;
    float* depth = nullptr;
    uint32_t* ids = nullptr;
    size_t bytes = 0;
};

struct axr_batch {
    axr_ctx* ctx;
    uint32_t nviews;
    std::vector<ViewDesc> views;
    ViewDesc* d_views = nullptr;
    uint32_t* d_face_bases = nullptr;
    uint32_t* d_tile_bases = nullptr;
    uint32_t* d_pix_bases = nullptr;
    uint32_t* d_tile_views = nullptr;
    uint32_t total_pixels = 0;
    bool clear = false;            // clear every view's framebuffer before drawing
    uint32_t clear_bgra = 0;
    float clear_depth = 0.0f;
    KArgs ka{};
    size_t bytes = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    uint64_t captured_gen = ~0ull;
    int captured_tile = 0;
    uint64_t clear_version = 0, captured_clear = ~0ull;
};

namespace {

int fail(axr_ctx* ctx, int code, const char* fmt, ...) {
    if (ctx) {
        char buf[512];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(buf, sizeof(buf), fmt, ap);
        va_end(ap);
        ctx->err = buf;
    }
    return code;
}

int hip_fail(axr_ctx* ctx, hipError_t e, const char* what) {
    return fail(ctx, AXR_ERR_HIP, "%s: %s", what, hipGetErrorString(e));
}

// Device allocation charged to the context's HBM budget.
int dev_alloc(axr_ctx* ctx, void** p, size_t bytes, bool charge = true) {
    *p = nullptr;
    if (bytes == 0) return AXR_OK;
    if (charge && ctx->budget && ctx->in_use + ctx->arena_bytes + bytes > ctx->budget)
        return fail(ctx, AXR_ERR_BUDGET, "allocation of %zu bytes exceeds the HBM budget (%llu in use of %llu)",
                    bytes, (unsigned long long)(ctx->in_use + ctx->arena_bytes), (unsigned long long)ctx->budget);
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) {
        *p = nullptr;
        return fail(ctx, AXR_ERR_OOM, "hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
    }
    return AXR_OK;
}

int upload(axr_ctx* ctx, void* dst, const void* src, size_t bytes) {
    if (!bytes) return AXR_OK;
    hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    return e == hipSuccess ? AXR_OK : hip_fail(ctx, e, "upload");
}

void free_buf(axr_ctx* ctx, DevBuf& b) {
    if (b.p) hipFree(b.p);
    ctx->arena_bytes -= b.bytes;
    b.p = nullptr;
    b.bytes = 0;
}

int grow(axr_ctx* ctx, DevBuf& b, size_t bytes) {
    if (b.bytes >= bytes) return AXR_OK;
    bytes = std::max(bytes, b.bytes + b.bytes / 4);
    hipError_t e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return hip_fail(ctx, e, "sync before arena growth");
    free_buf(ctx, b);
    if (ctx->budget && ctx->in_use + ctx->arena_bytes + bytes > ctx->budget)
        return fail(ctx, AXR_ERR_BUDGET, "frame arena of %zu more bytes exceeds the HBM budget", bytes);
    e = hipMalloc(&b.p, bytes);
    if (e != hipSuccess) { b.p = nullptr; return fail(ctx, AXR_ERR_OOM, "arena hipMalloc(%zu): %s", bytes, hipGetErrorString(e)); }
    b.bytes = bytes;
    ctx->arena_bytes += bytes;
    ++ctx->arena_gen;
    return AXR_OK;
}

// Size the frame arena for `faces` face slots and `tiles` bins; `need_*`
// carry high-water marks reported by the kernels after an overflow.
int ensure_arena(axr_ctx* ctx, uint64_t faces, uint64_t tiles, uint64_t need_clip_faces = 0,
                 uint64_t need_clip_recs = 0, uint64_t need_pairs = 0) {
    if (faces > 0x7FFFFFFFull || tiles > 0x7FFFFFFFull)
        return fail(ctx, AXR_ERR_INVALID, "draw too large (%llu faces, %llu bins)", (unsigned long long)faces,
                    (unsigned long long)tiles);
    Arena& a = ctx->arena;
    uint64_t clip_faces = std::max<uint64_t>({4096, faces / 16, need_clip_faces, a.cap_clip_faces});
    uint64_t clip_recs = std::max<uint64_t>({2 * clip_faces, need_clip_recs, a.cap_clip_recs});
    // one pair region per setup wave (setup runs 4 waves per 256 faces)
    uint64_t waves = std::max<uint64_t>(4 * ((faces + 255) / 256), a.cap_waves);
    uint64_t ovf = std::max<uint64_t>({faces / 4 + 16 * tiles + 4096, need_pairs, a.cap_ovf});
    uint64_t nbins = std::max<uint64_t>({2 * faces + 16 * tiles + 4096, need_pairs, a.cap_bins});
    uint64_t scan_blocks = (tiles + 4095) / 4096 + 1;
    int r;
    const uint64_t gen_start = ctx->arena_gen;
    if ((r = grow(ctx, ctx->counters, CNT_COUNT * 4))) return r;
    if ((r = grow(ctx, ctx->rec, std::max<uint64_t>(faces, 1) * sizeof(TriRec)))) return r;
    if ((r = grow(ctx, ctx->clip_faces, clip_faces * 4))) return r;
    if ((r = grow(ctx, ctx->clip_rec, clip_recs * sizeof(TriRec)))) return r;
    if ((r = grow(ctx, ctx->clip_verts, clip_recs * 3 * sizeof(Vertex56)))) return r;
    if ((r = grow(ctx, ctx->tile_count, (tiles + 1) * 4))) return r;
    if ((r = grow(ctx, ctx->tile_off, (tiles + 1) * 4))) return r;
    const uint32_t wave_cap = (uint32_t)ctx->cfg.bin_capacity;
    if ((r = grow(ctx, ctx->pairs, (waves * wave_cap + ovf) * 16))) return r;
    if ((r = grow(ctx, ctx->wave_pairs, waves * 4))) return r;
    if ((r = grow(ctx, ctx->bins, nbins * 4))) return r;
    if ((r = grow(ctx, ctx->active, (tiles + 1) * 4))) return r;
    if ((r = grow(ctx, ctx->scan_sums, scan_blocks * 8))) return r;
    a.counters = (uint32_t*)ctx->counters.p;
    a.rec = (TriRec*)ctx->rec.p;
    a.clip_faces = (uint32_t*)ctx->clip_faces.p;
    a.clip_rec = (TriRec*)ctx->clip_rec.p;
    a.clip_verts = (Vertex56*)ctx->clip_verts.p;
    a.tile_count = (uint32_t*)ctx->tile_count.p;
    a.tile_off = (uint32_t*)ctx->tile_off.p;
    a.pairs = (uint4*)ctx->pairs.p;
    a.wave_pairs = (uint32_t*)ctx->wave_pairs.p;
    a.bins = (uint32_t*)ctx->bins.p;
    a.active = (uint32_t*)ctx->active.p;
    a.scan_sums = (unsigned long long*)ctx->scan_sums.p;
    a.cap_clip_faces = (uint32_t)std::min<uint64_t>(ctx->clip_faces.bytes / 4, 0xFFFFFFFFull);
    a.cap_clip_recs = (uint32_t)std::min<uint64_t>(
        {ctx->clip_rec.bytes / sizeof(TriRec), ctx->clip_verts.bytes / (3 * sizeof(Vertex56)), 0x7FFFFFFFull});
    if (a.wave_cap != wave_cap) ++ctx->arena_gen;   // the pair layout changed: captured graphs are stale
    a.wave_cap = wave_cap;
    a.cap_waves = (uint32_t)std::min<uint64_t>(
        {ctx->wave_pairs.bytes / 4, 0x7FFFFFFFull / wave_cap, ctx->pairs.bytes / 16 / wave_cap});
    a.cap_ovf = (uint32_t)std::min<uint64_t>(ctx->pairs.bytes / 16 - (uint64_t)a.cap_waves * wave_cap, 0x7FFFFFFFull);
    a.cap_bins = (uint32_t)std::min<uint64_t>(ctx->bins.bytes / 4, 0x7FFFFFFFull);
    if (ctx->arena_gen != gen_start || ctx->tile_count.bytes / 4 < tiles + 1) ctx->arena_dirty = true;
    return AXR_OK;
}

// Zero the whole counter block (including the sticky error and the
// high-water marks).  Not used inside captured graphs.
int reset_counters(axr_ctx* ctx) {
    std::memset(ctx->host_counters, 0, CNT_COUNT * 4);
    hipError_t e = hipMemcpyAsync(ctx->counters.p, ctx->host_counters, CNT_COUNT * 4, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    return e == hipSuccess ? AXR_OK : hip_fail(ctx, e, "reset counters");
}

int read_counters(axr_ctx* ctx) {
    hipError_t e = hipMemcpyAsync(ctx->host_counters, ctx->counters.p, CNT_COUNT * 4, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    return e == hipSuccess ? AXR_OK : hip_fail(ctx, e, "read counters");
}

void harvest_events(axr_ctx* ctx) {
    while (ctx->ev_pending > 0) {
        int slot = (ctx->ev_head - ctx->ev_pending + kEventRing) % kEventRing;
        hipEvent_t* e = ctx->ev[slot];
        (void)hipEventSynchronize(e[5]);
        float total = 0;
        (void)hipEventElapsedTime(&total, e[0], e[5]);
        ctx->stats.frames += 1;
        ctx->stats.frame_ms_total += total;
        for (int k = 0; k < 5; ++k) {
            float ms = 0;
            (void)hipEventElapsedTime(&ms, e[k], e[k + 1]);
            ctx->stats.stage_ms_total[k] += ms;
        }
        --ctx->ev_pending;
    }
}

// Events for one frame: [0] before the frame, [1..5] after setup, clip,
// scan, scatter and raster.
hipEvent_t* next_events(axr_ctx* ctx) {
    if (!ctx->profiling) return nullptr;
    if (ctx->ev_pending == kEventRing) harvest_events(ctx);
    hipEvent_t* ev = ctx->ev[ctx->ev_head];
    ctx->ev_head = (ctx->ev_head + 1) % kEventRing;
    ++ctx->ev_pending;
    return ev;
}

// Zero the bin counts and per-frame counters if the arena is fresh or an
// overflow left them dirty.  Never inside a captured graph.
int prep_if_dirty(axr_ctx* ctx) {
    if (!ctx->arena_dirty) return AXR_OK;
    hipError_t e = launch_prep(ctx->arena, (uint32_t)(ctx->tile_count.bytes / 4), ctx->stream);
    if (e != hipSuccess) return hip_fail(ctx, e, "arena reset");
    ctx->arena_dirty = false;
    return AXR_OK;
}

int launch(axr_ctx* ctx, const KArgs& ka) {
    int r = prep_if_dirty(ctx);
    if (r) return r;
    hipEvent_t* ev = next_events(ctx);
    if (ev) (void)hipEventRecord(ev[0], ctx->stream);
    hipError_t e = launch_frame(ka, ctx->arena, ctx->stream, ev ? ev + 1 : nullptr);
    return e == hipSuccess ? AXR_OK : hip_fail(ctx, e, "kernel launch");
}

// After a synchronising read of the counters: AXR_OK, or grow the arena to
// the reported high-water marks and return AXR_ERR_OVERFLOW.
int check_overflow(axr_ctx* ctx, uint64_t faces, uint64_t tiles) {
    const uint32_t* c = ctx->host_counters;
    if (!c[CNT_ERROR]) return AXR_OK;
    uint64_t nf = c[CNT_NEED_CLIP_FACES], nr = c[CNT_NEED_CLIP_RECS], np = c[CNT_NEED_PAIRS];
    int r = ensure_arena(ctx, faces, tiles, nf ? nf + nf / 2 : 0, nr ? nr + nr / 2 : 0, np ? np + np / 4 : 0);
    if (r) return r;
    if ((r = reset_counters(ctx))) return r;
    ctx->arena_dirty = true;
    return fail(ctx, AXR_ERR_OVERFLOW, "frame arena overflow (clip faces %llu, clip records %llu, bin pairs %llu); "
                "arena grown, frames since the last sync were not drawn",
                (unsigned long long)nf, (unsigned long long)nr, (unsigned long long)np);
}

ViewDesc make_view(const axr_mesh* m, const axr_fb* fb, const axr_draw_params* p, int tile) {
    ViewDesc v{};
    v.pos = m->pos;
    v.verts = m->verts;
    v.idx = m->idx;
    v.face_colors = m->face_colors;
    v.groups = m->groups;
    v.materials = m->materials;
    v.ngroups = m->ngroups;
    v.nfaces = m->nfaces_sub;
    mat4 model, vp, viewport;
    std::memcpy(&model, p->model, 64);
    std::memcpy(&vp, p->view_proj, 64);
    std::memcpy(&viewport, p->viewport, 64);
    (void)viewport;   // the reference sets it as a uniform but no shader reads it
    // reference tiled_pipeline.cpp:148-155: mvp = viewProj * model
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
    v.color = fb->color;
    v.depth = fb->depth;
    v.ids = fb->ids;
    v.W = fb->w;
    v.H = fb->h;
    v.tiles_x = (fb->w + tile - 1) / tile;
    v.tiles_y = (fb->h + tile - 1) / tile;
    return v;
}

bool shader_ok(const axr_mesh* m, const axr_draw_params* p, axr_ctx* ctx) {
    if (p->shader < 0 || p->shader >= SHADER_COUNT) { fail(ctx, AXR_ERR_INVALID, "unknown shader %d", p->shader); return false; }
    return true;
}

}  // namespace

extern "C" {

const char* axr_version(void) { return "axiomr-mi355x 0.1 (gfx950)"; }

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
    hipError_t e = hipSetDevice(device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipHostMalloc((void**)&ctx->host_counters, CNT_COUNT * 4, hipHostMallocDefault);
    if (e != hipSuccess) {
        if (ctx->stream) hipStreamDestroy(ctx->stream);
        delete ctx;
        return AXR_ERR_HIP;
    }
    int r = ensure_arena(ctx, 1024, 1024);
    if (r == AXR_OK) r = reset_counters(ctx);
    if (r != AXR_OK) { axr_ctx_destroy(ctx); return r; }
    *out = ctx;
    return AXR_OK;
}

int axr_ctx_destroy(axr_ctx* ctx) {
    if (!ctx) return AXR_ERR_INVALID;
    hipSetDevice(ctx->device);
    if (ctx->stream) hipStreamSynchronize(ctx->stream);
    DevBuf* bufs[] = {&ctx->counters, &ctx->rec, &ctx->clip_faces, &ctx->clip_rec, &ctx->clip_verts,
                      &ctx->tile_count, &ctx->tile_off, &ctx->pairs, &ctx->wave_pairs, &ctx->bins, &ctx->active,
                      &ctx->scan_sums};
    for (DevBuf* b : bufs) free_buf(ctx, *b);
    for (auto& ring : ctx->ev)
        for (auto& e : ring)
            if (e) hipEventDestroy(e);
    if (ctx->host_counters) hipHostFree(ctx->host_counters);
    if (ctx->stream) hipStreamDestroy(ctx->stream);
    delete ctx;
    return AXR_OK;
}

const char* axr_last_error(const axr_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int axr_ctx_set_config(axr_ctx* ctx, const axr_config* cfg) {
    if (!ctx || !cfg) return AXR_ERR_INVALID;
    if (cfg->tile_size != 8 && cfg->tile_size != 16 && cfg->tile_size != 32)
        return fail(ctx, AXR_ERR_INVALID, "tile_size must be 8, 16 or 32 (got %d)", cfg->tile_size);
    if (cfg->raster_threads != 64 && cfg->raster_threads != 128 && cfg->raster_threads != 256)
        return fail(ctx, AXR_ERR_INVALID, "raster_threads must be 64, 128 or 256 (got %d)", cfg->raster_threads);
    if (cfg->bin_capacity < 64 || cfg->bin_capacity > 1024)
        return fail(ctx, AXR_ERR_INVALID, "bin_capacity must be in [64, 1024] (got %d)", cfg->bin_capacity);
    if (cfg->small_footprint < 1 || cfg->small_footprint > 1024)
        return fail(ctx, AXR_ERR_INVALID, "small_footprint must be in [1, 1024] (got %d)", cfg->small_footprint);
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
    hipSetDevice(ctx->device);
    if (enable && !ctx->ev[0][0]) {
        for (auto& ring : ctx->ev)
            for (auto& e : ring)
                if (hipEventCreate(&e) != hipSuccess) return fail(ctx, AXR_ERR_HIP, "hipEventCreate");
    }
    if (!enable) harvest_events(ctx);
    ctx->profiling = enable != 0;
    ctx->stats = axr_kernel_stats{};
    return AXR_OK;
}

int axr_ctx_get_stats(axr_ctx* ctx, axr_kernel_stats* out) {
    if (!ctx || !out) return AXR_ERR_INVALID;
    hipSetDevice(ctx->device);
    harvest_events(ctx);
    *out = ctx->stats;
    return AXR_OK;
}

int axr_ctx_memory(const axr_ctx* ctx, uint64_t* in_use, uint64_t* arena) {
    if (!ctx) return AXR_ERR_INVALID;
    if (in_use) *in_use = ctx->in_use;
    if (arena) *arena = ctx->arena_bytes;
    return AXR_OK;
}

int axr_sync(axr_ctx* ctx) {
    if (!ctx) return AXR_ERR_INVALID;
    hipSetDevice(ctx->device);
    hipError_t e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return hip_fail(ctx, e, "hipStreamSynchronize");
    int r = read_counters(ctx);
    if (r) return r;
    if (ctx->host_counters[CNT_ERROR]) {
        // sizes of the last failing draw are what the kernels reported; keep
        // the current face/bin capacity as the floor
        uint64_t faces = ctx->rec.bytes / sizeof(TriRec);
        uint64_t tiles = ctx->tile_count.bytes / 4 - 1;
        return check_overflow(ctx, faces, tiles);
    }
    return AXR_OK;
}

// ---- resources -------------------------------------------------------------

int axr_texture_create(axr_ctx* ctx, const uint8_t* rgba, int w, int h, int filter, axr_texture** out) {
    if (!ctx || !out) return AXR_ERR_INVALID;
    *out = nullptr;
    if (!rgba || w <= 0 || h <= 0 || (uint64_t)w * (uint64_t)h > (1ull << 30))
        return fail(ctx, AXR_ERR_INVALID, "texture: bad size %dx%d", w, h);
    hipSetDevice(ctx->device);
    axr_texture* t = new (std::nothrow) axr_texture();
    if (!t) return AXR_ERR_OOM;
    t->ctx = ctx;
    t->bytes = (size_t)w * h * 4;
    void* p;
    int r = dev_alloc(ctx, &p, t->bytes);
    if (r) { delete t; return r; }
    if ((r = upload(ctx, p, rgba, t->bytes))) { hipFree(p); delete t; return r; }
    t->dev = DevTexture{(const uint32_t*)p, w, h, filter ? 1 : 0, 0};
    ctx->in_use += t->bytes;
    *out = t;
    return AXR_OK;
}

int axr_texture_destroy(axr_texture* t) {
    if (!t) return AXR_ERR_INVALID;
    hipSetDevice(t->ctx->device);
    hipStreamSynchronize(t->ctx->stream);
    hipFree((void*)t->dev.texels);
    t->ctx->in_use -= t->bytes;
    delete t;
    return AXR_OK;
}

int axr_mesh_create(axr_ctx* ctx, const float* verts56, uint64_t nverts, const uint32_t* idx, uint64_t nfaces,
                    const axr_group* groups, uint32_t ngroups, const axr_material* materials, uint32_t nmaterials,
                    const uint32_t* face_colors, axr_mesh** out) {
    if (!ctx || !out) return AXR_ERR_INVALID;
    *out = nullptr;
    if (!verts56 || !idx || nverts == 0 || nfaces == 0 || nverts > 0xFFFFFFFFull || nfaces > 0x0FFFFFFFull)
        return fail(ctx, AXR_ERR_INVALID, "mesh: bad sizes (%llu vertices, %llu faces)", (unsigned long long)nverts,
                    (unsigned long long)nfaces);
    if (!groups || ngroups == 0 || !materials || nmaterials == 0 || nmaterials > REC_MATERIAL_MASK)
        return fail(ctx, AXR_ERR_INVALID, "mesh: needs at least one group and one material");
    for (uint64_t i = 0; i < 3 * nfaces; ++i)
        if (idx[i] >= nverts) return fail(ctx, AXR_ERR_INVALID, "mesh: index %u at %llu out of range", idx[i], (unsigned long long)i);
    std::vector<DevGroup> dg(ngroups);
    uint64_t rank = 0;
    for (uint32_t g = 0; g < ngroups; ++g) {
        const axr_group& s = groups[g];
        if ((uint64_t)s.first_face + s.face_count > nfaces)
            return fail(ctx, AXR_ERR_INVALID, "mesh: group %u covers faces past the end", g);
        // reference Mesh::getMaterial uses .at() and throws (include/mesh.hpp:80-82)
        if (s.material >= nmaterials) return fail(ctx, AXR_ERR_INVALID, "mesh: group %u names material %u of %u", g, s.material, nmaterials);
        dg[g] = DevGroup{s.first_face, s.face_count, (uint32_t)rank, s.material};
        rank += s.face_count;
    }
    if (rank > 0x1FFFFFFFull) return fail(ctx, AXR_ERR_INVALID, "mesh: too many faces in submission order");
    std::vector<DevMaterial> dm(nmaterials);
    for (uint32_t i = 0; i < nmaterials; ++i) {
        const axr_material& s = materials[i];
        auto tex = [](const axr_texture* t) { return t ? t->dev : DevTexture{nullptr, 0, 0, 0, 0}; };
        dm[i] = DevMaterial{tex(s.diffuse), tex(s.metallic), tex(s.bump), tex(s.roughness), tex(s.ao), s.specular_exponent, {0, 0, 0}};
    }
    hipSetDevice(ctx->device);
    axr_mesh* m = new (std::nothrow) axr_mesh();
    if (!m) return AXR_ERR_OOM;
    m->ctx = ctx;
    m->nverts = nverts;
    m->nfaces = nfaces;
    m->ngroups = ngroups;
    m->nmaterials = nmaterials;
    m->nfaces_sub = (uint32_t)rank;
    std::vector<float> pos(3 * nverts);
    for (uint64_t i = 0; i < nverts; ++i)
        for (int k = 0; k < 3; ++k) pos[3 * i + k] = verts56[14 * i + k];
    const size_t sz_pos = pos.size() * 4, sz_v = nverts * 56, sz_i = nfaces * 12;
    const size_t sz_c = face_colors ? nfaces * 4 : 0, sz_g = ngroups * sizeof(DevGroup), sz_m = nmaterials * sizeof(DevMaterial);
    int r = AXR_OK;
    auto cleanup = [&]() {
        void* ps[] = {m->pos, m->verts, m->idx, m->face_colors, m->groups, m->materials};
        for (void* p : ps) if (p) hipFree(p);
        delete m;
    };
    if ((r = dev_alloc(ctx, (void**)&m->pos, sz_pos)) || (r = dev_alloc(ctx, (void**)&m->verts, sz_v)) ||
        (r = dev_alloc(ctx, (void**)&m->idx, sz_i)) || (r = dev_alloc(ctx, (void**)&m->face_colors, sz_c)) ||
        (r = dev_alloc(ctx, (void**)&m->groups, sz_g)) || (r = dev_alloc(ctx, (void**)&m->materials, sz_m))) {
        cleanup();
        return r;
    }
    if ((r = upload(ctx, m->pos, pos.data(), sz_pos)) || (r = upload(ctx, m->verts, verts56, sz_v)) ||
        (r = upload(ctx, m->idx, idx, sz_i)) || (r = upload(ctx, m->face_colors, face_colors, sz_c)) ||
        (r = upload(ctx, m->groups, dg.data(), sz_g)) || (r = upload(ctx, m->materials, dm.data(), sz_m))) {
        cleanup();
        return r;
    }
    m->bytes = sz_pos + sz_v + sz_i + sz_c + sz_g + sz_m;
    ctx->in_use += m->bytes;
    *out = m;
    return AXR_OK;
}

int axr_mesh_destroy(axr_mesh* m) {
    if (!m) return AXR_ERR_INVALID;
    hipSetDevice(m->ctx->device);
    hipStreamSynchronize(m->ctx->stream);
    void* ps[] = {m->pos, m->verts, m->idx, m->face_colors, m->groups, m->materials};
    for (void* p : ps) if (p) hipFree(p);
    m->ctx->in_use -= m->bytes;
    delete m;
    return AXR_OK;
}

int axr_fb_create(axr_ctx* ctx, int w, int h, int with_ids, axr_fb** out) {
    if (!ctx || !out) return AXR_ERR_INVALID;
    *out = nullptr;
    if (w <= 0 || h <= 0 || w > 32768 || h > 32768) return fail(ctx, AXR_ERR_INVALID, "framebuffer: bad size %dx%d", w, h);
    hipSetDevice(ctx->device);
    axr_fb* fb = new (std::nothrow) axr_fb();
    if (!fb) return AXR_ERR_OOM;
    fb->ctx = ctx;
    fb->w = w;
    fb->h = h;
    const size_t n = (size_t)w * h;
    int r;
    if ((r = dev_alloc(ctx, (void**)&fb->color, n * 4)) || (r = dev_alloc(ctx, (void**)&fb->depth, n * 4)) ||
        (with_ids && (r = dev_alloc(ctx, (void**)&fb->ids, n * 4)))) {
        if (fb->color) hipFree(fb->color);
        if (fb->depth) hipFree(fb->depth);
        if (fb->ids) hipFree(fb->ids);
        delete fb;
        return r;
    }
    fb->bytes = n * (with_ids ? 12 : 8);
    ctx->in_use += fb->bytes;
    // reference Framebuffer ctor: colour 0, depth +inf (src/framebuffer.cpp:6-14)
    hipError_t e = launch_clear(fb->color, fb->depth, fb->ids, n, 0u, __builtin_huge_valf(), ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) { axr_fb_destroy(fb); return hip_fail(ctx, e, "framebuffer init"); }
    *out = fb;
    return AXR_OK;
}

int axr_fb_destroy(axr_fb* fb) {
    if (!fb) return AXR_ERR_INVALID;
    hipSetDevice(fb->ctx->device);
    hipStreamSynchronize(fb->ctx->stream);
    if (fb->color) hipFree(fb->color);
    if (fb->depth) hipFree(fb->depth);
    if (fb->ids) hipFree(fb->ids);
    fb->ctx->in_use -= fb->bytes;
    delete fb;
    return AXR_OK;
}

int axr_fb_clear(axr_fb* fb, uint32_t bgra, float depth) {
    if (!fb) return AXR_ERR_INVALID;
    hipSetDevice(fb->ctx->device);
    hipError_t e = launch_clear(fb->color, fb->depth, fb->ids, (size_t)fb->w * fb->h, bgra, depth, fb->ctx->stream);
    return e == hipSuccess ? AXR_OK : hip_fail(fb->ctx, e, "framebuffer clear");
}

int axr_fb_read(axr_fb* fb, uint8_t* bgra, float* depth, uint32_t* ids) {
    if (!fb) return AXR_ERR_INVALID;
    axr_ctx* ctx = fb->ctx;
    hipSetDevice(ctx->device);
    const size_t n = (size_t)fb->w * fb->h;
    hipError_t e = hipStreamSynchronize(ctx->stream);
    if (e == hipSuccess && bgra) e = hipMemcpy(bgra, fb->color, n * 4, hipMemcpyDeviceToHost);
    if (e == hipSuccess && depth) e = hipMemcpy(depth, fb->depth, n * 4, hipMemcpyDeviceToHost);
    if (e == hipSuccess && ids) {
        if (!fb->ids) return fail(ctx, AXR_ERR_INVALID, "framebuffer has no id plane");
        e = hipMemcpy(ids, fb->ids, n * 4, hipMemcpyDeviceToHost);
    }
    return e == hipSuccess ? AXR_OK : hip_fail(ctx, e, "framebuffer read");
}

// ---- drawMesh ----------------------------------------------------------------

int axr_draw(axr_ctx* ctx, axr_fb* fb, const axr_mesh* mesh, const axr_draw_params* p) {
    if (!ctx) return AXR_ERR_INVALID;
    // reference drawMesh returns silently without a shader/camera/framebuffer
    // (tiled_pipeline.cpp:146); over the C ABI a missing argument is an error.
    if (!fb || !mesh || !p) return fail(ctx, AXR_ERR_INVALID, "axr_draw: null framebuffer, mesh or params");
    if (fb->ctx != ctx || mesh->ctx != ctx) return fail(ctx, AXR_ERR_INVALID, "axr_draw: handle from another context");
    if (!shader_ok(mesh, p, ctx)) return AXR_ERR_INVALID;
    hipSetDevice(ctx->device);
    const int tile = ctx->cfg.tile_size;
    KArgs ka{};
    ka.table = nullptr;
    ka.nviews = 1;
    ka.one = make_view(mesh, fb, p, tile);
    ka.one.face_base = 0;
    ka.one.tile_base = 0;
    ka.total_faces = mesh->nfaces_sub;
    ka.total_tiles = (uint32_t)(ka.one.tiles_x * ka.one.tiles_y);
    ka.tile_size = tile;
    ka.raster_threads = ctx->cfg.raster_threads;
    ka.small_footprint = ctx->cfg.small_footprint;
    int r = ensure_arena(ctx, ka.total_faces, ka.total_tiles);
    if (r) return r;
    const bool async = (p->flags & AXR_DRAW_ASYNC) != 0;
    for (int attempt = 0; attempt < 8; ++attempt) {
        if ((r = launch(ctx, ka))) return r;
        if (async) return AXR_OK;
        hipError_t e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) return hip_fail(ctx, e, "draw");
        if ((r = read_counters(ctx))) return r;
        if (!ctx->host_counters[CNT_ERROR]) return AXR_OK;
        r = check_overflow(ctx, ka.total_faces, ka.total_tiles);
        if (r != AXR_ERR_OVERFLOW) return r;
        // the raster stage was skipped, so the framebuffer is untouched: redraw
    }
    return fail(ctx, AXR_ERR_OVERFLOW, "axr_draw: frame arena kept overflowing");
}

// ---- batched form --------------------------------------------------------------

int axr_batch_create(axr_ctx* ctx, uint32_t nviews, axr_fb* const* fbs, const axr_mesh* const* meshes,
                     const axr_draw_params* params, axr_batch** out) {
    if (!ctx || !out) return AXR_ERR_INVALID;
    *out = nullptr;
    if (nviews == 0 || !fbs || !meshes || !params) return fail(ctx, AXR_ERR_INVALID, "batch: empty or null arrays");
    hipSetDevice(ctx->device);
    axr_batch* b = new (std::nothrow) axr_batch();
    if (!b) return AXR_ERR_OOM;
    b->ctx = ctx;
    b->nviews = nviews;
    b->views.resize(nviews);
    std::vector<uint32_t> fbase(nviews + 1), tbase(nviews + 1), pbase(nviews + 1);
    uint64_t faces = 0, tiles = 0, pixels = 0;
    const int tile = ctx->cfg.tile_size;
    for (uint32_t i = 0; i < nviews; ++i) {
        if (!fbs[i] || !meshes[i] || fbs[i]->ctx != ctx || meshes[i]->ctx != ctx || !shader_ok(meshes[i], &params[i], ctx)) {
            delete b;
            return fail(ctx, AXR_ERR_INVALID, "batch: view %u has a bad handle or shader", i);
        }
        ViewDesc v = make_view(meshes[i], fbs[i], &params[i], tile);
        v.face_base = (uint32_t)faces;
        v.tile_base = (uint32_t)tiles;
        fbase[i] = (uint32_t)faces;
        tbase[i] = (uint32_t)tiles;
        pbase[i] = (uint32_t)pixels;
        faces += v.nfaces;
        tiles += (uint64_t)v.tiles_x * v.tiles_y;
        pixels += (uint64_t)v.W * v.H;
        if (faces > 0x7FFFFFFFull || tiles > 0x7FFFFFFFull || pixels > 0xFFFFFFFFull) {
            delete b;
            return fail(ctx, AXR_ERR_INVALID, "batch too large");
        }
        b->views[i] = v;
    }
    fbase[nviews] = (uint32_t)faces;
    tbase[nviews] = (uint32_t)tiles;
    pbase[nviews] = (uint32_t)pixels;
    b->total_pixels = (uint32_t)pixels;
    std::vector<uint32_t> tview(tiles);
    for (uint32_t i = 0; i < nviews; ++i)
        for (uint32_t t = tbase[i]; t < tbase[i + 1]; ++t) tview[t] = i;
    int r;
    const size_t sz_v = nviews * sizeof(ViewDesc), sz_b = (nviews + 1) * 4;
    if ((r = dev_alloc(ctx, (void**)&b->d_views, sz_v)) || (r = dev_alloc(ctx, (void**)&b->d_face_bases, sz_b)) ||
        (r = dev_alloc(ctx, (void**)&b->d_tile_bases, sz_b)) || (r = dev_alloc(ctx, (void**)&b->d_pix_bases, sz_b)) ||
        (r = upload(ctx, b->d_views, b->views.data(), sz_v)) || (r = upload(ctx, b->d_face_bases, fbase.data(), sz_b)) ||
        (r = upload(ctx, b->d_tile_bases, tbase.data(), sz_b)) || (r = upload(ctx, b->d_pix_bases, pbase.data(), sz_b)) ||
        (r = dev_alloc(ctx, (void**)&b->d_tile_views, tiles * 4)) ||
        (r = upload(ctx, b->d_tile_views, tview.data(), tiles * 4))) {
        axr_batch_destroy(b);
        return r;
    }
    b->bytes = sz_v + 3 * sz_b + tiles * 4;
    ctx->in_use += b->bytes;
    KArgs& ka = b->ka;
    ka.table = b->d_views;
    ka.face_bases = b->d_face_bases;
    ka.tile_bases = b->d_tile_bases;
    ka.tile_views = b->d_tile_views;
    ka.nviews = nviews;
    ka.total_faces = (uint32_t)faces;
    ka.total_tiles = (uint32_t)tiles;
    ka.tile_size = tile;
    ka.raster_threads = ctx->cfg.raster_threads;
    ka.small_footprint = ctx->cfg.small_footprint;
    ka.one = b->views[0];
    if ((r = ensure_arena(ctx, faces, tiles))) { axr_batch_destroy(b); return r; }
    *out = b;
    return AXR_OK;
}

int axr_batch_launch(axr_batch* b) {
    if (!b) return AXR_ERR_INVALID;
    axr_ctx* ctx = b->ctx;
    hipSetDevice(ctx->device);
    int r = ensure_arena(ctx, b->ka.total_faces, b->ka.total_tiles);
    if (r) return r;
    if (ctx->profiling) {   // events need the eager path
        if (b->clear) {
            hipError_t e = launch_clear_views(b->d_views, b->d_pix_bases, b->nviews, b->total_pixels, b->clear_bgra,
                                              b->clear_depth, ctx->stream);
            if (e != hipSuccess) return hip_fail(ctx, e, "batch clear");
        }
        return launch(ctx, b->ka);
    }
    if ((r = prep_if_dirty(ctx))) return r;
    if (!b->exec || b->captured_gen != ctx->arena_gen || b->captured_tile != b->ka.tile_size ||
        b->captured_clear != b->clear_version) {
        if (b->exec) { hipGraphExecDestroy(b->exec); b->exec = nullptr; }
        if (b->graph) { hipGraphDestroy(b->graph); b->graph = nullptr; }
        hipError_t e = hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal);
        if (e != hipSuccess) return hip_fail(ctx, e, "hipStreamBeginCapture");
        hipError_t le = hipSuccess;
        if (b->clear)
            le = launch_clear_views(b->d_views, b->d_pix_bases, b->nviews, b->total_pixels, b->clear_bgra,
                                    b->clear_depth, ctx->stream);
        if (le == hipSuccess) le = launch_frame(b->ka, ctx->arena, ctx->stream, nullptr);
        e = hipStreamEndCapture(ctx->stream, &b->graph);
        if (le != hipSuccess) return hip_fail(ctx, le, "batch capture launch");
        if (e != hipSuccess) return hip_fail(ctx, e, "hipStreamEndCapture");
        e = hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0);
        if (e != hipSuccess) return hip_fail(ctx, e, "hipGraphInstantiate");
        b->captured_gen = ctx->arena_gen;
        b->captured_tile = b->ka.tile_size;
        b->captured_clear = b->clear_version;
    }
    hipError_t e = hipGraphLaunch(b->exec, ctx->stream);
    return e == hipSuccess ? AXR_OK : hip_fail(ctx, e, "hipGraphLaunch");
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
    hipSetDevice(ctx->device);
    hipStreamSynchronize(ctx->stream);
    if (b->exec) hipGraphExecDestroy(b->exec);
    if (b->graph) hipGraphDestroy(b->graph);
    if (b->d_views) hipFree(b->d_views);
    if (b->d_face_bases) hipFree(b->d_face_bases);
    if (b->d_tile_bases) hipFree(b->d_tile_bases);
    if (b->d_pix_bases) hipFree(b->d_pix_bases);
    if (b->d_tile_views) hipFree(b->d_tile_views);
    ctx->in_use -= b->bytes;
    delete b;
    return AXR_OK;
}

}  // extern "C"
