// noalloc.hip - axr_draw_noalloc: drawMesh that allocates nothing.
//
// Some callers cannot allocate device memory while they draw (code under
// HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
// axr_draw keeps its working memory in frame arenas that grow with the
// frame; this entry point instead carves the frame's working memory out of
// one file-scope device array, `scratch`, of a fixed size per GPU (held by
// every process that loads the library, until it exits).  It runs the
// visibility-mode kernels of axr_draw on the context's own stream, so its
// pixels are those of axr_draw.  Frames that need more than the array holds
// are rejected before any work is queued; a frame whose clipper or
// large-triangle queues overflow their share is not drawn and reports
// AXR_ERR_OVERFLOW (at the call when synchronous, else at axr_sync).
//
// Every call on one GPU uses that GPU's one array: calls on one context are
// ordered by its stream, but calls from different contexts or threads that
// overlap on one GPU race.  The array keeps the last call's contents, so
// each call first clears what the kernels expect to start at zero (k_prep:
// per-frame counters, bin flags, the per-pixel keys).
#include <cstring>
#include <mutex>
#include "internal.hpp"

namespace axr {

constexpr uint64_t NOALLOC_SCRATCH_BYTES = 64ull << 20;

// The frame's working memory on each GPU.  In host code `scratch` names a
// host-side stand-in, not device memory: every host use goes through the
// device address hipGetSymbolAddress returns (scratch_ptr).
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

namespace {

__global__ void k_zero_counters(uint32_t* counters) {
    if (threadIdx.x < CNT_COUNT) counters[threadIdx.x] = 0u;
}

constexpr int kMaxDevices = 64;
std::mutex g_mu;
unsigned char* g_ptr[kMaxDevices] = {};
bool g_fresh[kMaxDevices] = {};   // the counters have not been zeroed yet

// The device address of `scratch` on the current device (looked up once per
// device, with that device current).
int scratch_ptr(ErrSink& err, int device, unsigned char** out, bool* fresh) {
    if (device < 0 || device >= kMaxDevices) return fail(err, AXR_ERR_INVALID, "device %d out of range", device);
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_ptr[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) return hip_fail(err, e == hipSuccess ? hipErrorInvalidSymbol : e, "scratch address");
        g_ptr[device] = (unsigned char*)p;
        g_fresh[device] = true;
    }
    *out = g_ptr[device];
    *fresh = g_fresh[device];
    g_fresh[device] = false;
    return AXR_OK;
}

constexpr uint64_t up256(uint64_t x) { return (x + 255) & ~255ull; }

// Lays out a visibility-mode frame arena of `faces` face slots, `tiles`
// bins and `pixels` pixels (plus colour slots when `color_slots`) at base;
// returns the bytes it needs (0 if the sizes overflow the index types).
uint64_t carve(unsigned char* base, uint64_t faces, uint64_t tiles, uint64_t pixels, bool color_slots, Arena& a) {
    a = Arena{};
    const uint64_t clip_faces = std::max<uint64_t>(1024, faces / 64);
    const uint64_t clip_recs = 2 * clip_faces;
    const uint64_t big_recs = std::max<uint64_t>(1024, faces / 32);
    const uint64_t big_items = std::max<uint64_t>(16384, faces / 4 + pixels / 32);
    if (faces > MAX_BATCH_FACES || tiles > 0x7FFFFFFFull || pixels > 0xFFFFFFFFull || big_items > 0x7FFFFFFFull)
        return 0;
    uint64_t off = 0;
    auto take = [&](uint64_t bytes) {
        unsigned char* p = base ? base + off : nullptr;
        off += up256(bytes);
        return p;
    };
    a.counters = (uint32_t*)take(CNT_COUNT * 4);   // first, at a fixed offset: the error word is sticky
    a.vkeys = (unsigned long long*)take(pixels * 8);
    a.tile_count = (uint32_t*)take((tiles + 1) * 4);
    a.clip_base = (uint32_t*)take(std::max<uint64_t>(faces, 1) * 4);
    a.rec = (TriRec*)a.clip_base;
    a.clip_faces = (uint32_t*)take(clip_faces * 4);
    a.clip_rec = (TriRec*)take(clip_recs * sizeof(TriRec));
    a.clip_verts = (Vertex56*)take(clip_recs * 3 * sizeof(Vertex56));
    a.big_rec = (TriRec*)take(big_recs * sizeof(TriRec));
    a.big_start = (uint32_t*)take(big_recs * 4);
    a.vcol = color_slots ? (uint32_t*)take(pixels * 4) : nullptr;
    a.cap_clip_faces = (uint32_t)clip_faces;
    a.cap_clip_recs = (uint32_t)clip_recs;
    a.cap_vkeys = (uint32_t)pixels;
    a.cap_big_recs = (uint32_t)big_recs;
    a.cap_big_items = (uint32_t)big_items;
    a.wave_cap = 64;   // binned-mode fields stay empty: this path never bins
    return off;
}

}  // namespace

// The overflow an asynchronous call left in the scratch counters of the
// context's device: AXR_OK, or AXR_ERR_OVERFLOW after zeroing them.
// Called by axr_sync after the context's streams are idle.
int noalloc_settle(axr_ctx* ctx) {
    if (!ctx->noalloc_pending) return AXR_OK;
    ctx->noalloc_pending = false;
    unsigned char* base;
    bool fresh;
    int r = scratch_ptr(ctx->err, ctx->device, &base, &fresh);
    if (r) return r;
    uint32_t* counters = (uint32_t*)base;
    hipStream_t st = ctx->main.stream;
    hipError_t e = hipMemcpyAsync(ctx->noalloc_counters, counters, CNT_COUNT * 4, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return hip_fail(ctx->err, e, "scratch counters");
    if (!ctx->noalloc_counters[CNT_ERROR]) return AXR_OK;
    k_zero_counters<<<1, 64, 0, st>>>(counters);
    e = hipStreamSynchronize(st);
    if (e != hipSuccess) return hip_fail(ctx->err, e, "scratch counters");
    return fail(ctx->err, AXR_ERR_OVERFLOW,
                "axr_draw_noalloc: a frame needed more clipper or large-triangle room than its scratch share; "
                "frames since the last sync were not drawn");
}

}  // namespace axr

using namespace axr;

extern "C" int axr_draw_noalloc(axr_ctx* ctx, axr_fb* fb, const axr_mesh* mesh, const axr_draw_params* p) {
    if (!ctx) return AXR_ERR_INVALID;
    if (!fb || !mesh || !p) return fail(ctx->err, AXR_ERR_INVALID, "axr_draw_noalloc: null framebuffer, mesh or params");
    if (fb->ctx != ctx || mesh->ctx != ctx)
        return fail(ctx->err, AXR_ERR_INVALID, "axr_draw_noalloc: handle from another context");
    if (!shader_ok(p)) return fail(ctx->err, AXR_ERR_INVALID, "unknown shader %d", p->shader);
    if (axr_check_view(mesh->nfaces_sub, fb->w, fb->h))
        return fail(ctx->err, AXR_ERR_INVALID, "axr_draw_noalloc: view too large (%u faces, %dx%d)", mesh->nfaces_sub,
                    fb->w, fb->h);
    const int tile = ctx->cfg.tile_size;
    KArgs ka{};
    ka.nviews = 1;
    ka.one = make_view(mesh, fb->color, fb->depth, fb->ids, fb->w, fb->h, p, tile);
    ka.total_faces = mesh->nfaces_sub;
    ka.total_tiles = (uint32_t)(ka.one.tiles_x * ka.one.tiles_y);
    ka.tile_size = tile;
    ka.raster_threads = ctx->cfg.raster_threads;
    ka.small_footprint = ctx->cfg.small_footprint;
    ka.mode = RASTER_VISIBILITY;
    ka.total_pixels = (uint32_t)((uint64_t)fb->w * (uint64_t)fb->h);
    ka.shader = p->shader;
    const bool async = (p->flags & AXR_DRAW_ASYNC) != 0;
    if (fb->clear_pending) {
        // the resolve writes the recorded clear (as axr_draw does)
        ka.clear = 1;
        ka.clear_bgra = fb->clear_bgra;
        ka.clear_depth = fb->clear_depth;
        ka.two_pass = async;
    }
    const bool color_slots = ka.two_pass && fb->ids;
    Arena a;
    const uint64_t need = carve(nullptr, ka.total_faces, ka.total_tiles, ka.total_pixels, color_slots, a);
    if (need == 0 || need > NOALLOC_SCRATCH_BYTES)
        return fail(ctx->err, AXR_ERR_INVALID,
                    "axr_draw_noalloc: a %dx%d frame of %u faces needs %llu bytes of scratch; the limit is %llu",
                    fb->w, fb->h, mesh->nfaces_sub, (unsigned long long)need,
                    (unsigned long long)NOALLOC_SCRATCH_BYTES);
    (void)hipSetDevice(ctx->device);
    unsigned char* base;
    bool fresh;
    int r = scratch_ptr(ctx->err, ctx->device, &base, &fresh);
    if (r) return r;
    carve(base, ka.total_faces, ka.total_tiles, ka.total_pixels, color_slots, a);
    // after every earlier call on this context, on its own stream
    if ((r = ctx_join(ctx))) return r;
    hipStream_t st = ctx->main.stream;
    if (fresh) k_zero_counters<<<1, 64, 0, st>>>(a.counters);
    hipError_t e = launch_prep(a, ka.total_tiles, st);
    if (e == hipSuccess) e = launch_frame(ka, a, st, nullptr);
    if (e != hipSuccess) return hip_fail(ctx->err, e, "axr_draw_noalloc launch");
    fb->clear_pending = false;   // written by this launch, even by a frame that overflows
    ctx->noalloc_pending = true;
    if (async) return AXR_OK;
    return noalloc_settle(ctx);
}
