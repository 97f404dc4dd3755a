// internal.hpp - host-side objects behind the C ABI handles, shared by the
// context/draw/batch code (capi.hip) and the render server (serve.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/axiomr.h"
#include "axr_types.hpp"

namespace axr {

// What a frame's framebuffer pass waits for first: an event, or an 8-byte
// signal reaching at least `value` (stream memory wait, about a third of the
// host time of an event record + wait); nothing when both are empty.
struct FbWait {
    hipEvent_t ev = nullptr;
    uint64_t* sig = nullptr;
    uint64_t value = 0;
};

// Kernel timing of one frame: stage k's first kernel carries start[k] and
// its last kernel stop[k], attached to the dispatches themselves
// (hipExtLaunchKernel), so stop - start is the stage's kernel time as the
// command processor sees it (the figure a kernel trace reports), without
// the launch gap before it.  `used` has bit k set for every stage launched.
constexpr int kStages = 6;
struct StageEvents {
    hipEvent_t start[kStages] = {};
    hipEvent_t stop[kStages] = {};
    uint32_t used = 0;
};

// wait (optional): what the stream waits for before the pass that writes
// framebuffers (k_resolve / k_finish / k_raster); the passes before it only
// touch the frame arena and the meshes.  ev (optional): kernel timing.
hipError_t launch_frame(const KArgs& ka, const Arena& ar, hipStream_t st, StageEvents* ev, const FbWait& wait = {});
hipError_t launch_clear(uint32_t* color, float* depth, uint32_t* ids, size_t n, uint32_t bgra, float d, hipStream_t st);
hipError_t launch_prep(const Arena& ar, uint32_t ntiles, hipStream_t st);
hipError_t launch_clear_views(const ViewDesc* views, const uint32_t* pix_bases, uint32_t nviews, uint32_t total,
                              uint32_t bgra, float d, hipStream_t st);
// Bytes of a shader kind's vertex-stage output per vertex (0: none).
uint32_t vary_bytes(int kind);

struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
};

// Where a failing call leaves its message.  The context has one (read by
// axr_last_error); the server's scheduler thread has its own.
struct ErrSink {
    std::string msg;
};

// Capacities an overflowing frame reported (0: no requirement).
struct ArenaNeeds {
    uint64_t clip_faces = 0, clip_recs = 0, pairs = 0, big_recs = 0, big_items = 0;
};

// The per-frame working memory of the binning pipeline plus the stream its
// kernels run on.  A context owns one; each lane of a render server owns one,
// so batches on different lanes overlap on the device.
struct FrameArena {
    hipStream_t stream = nullptr;
    DevBuf counters, rec, clip_faces, clip_rec, clip_verts, tile_count, tile_off, pairs, wave_pairs, bins, active,
        scan_sums, vkeys, big_rec, big_start, vcol;
    Arena a{};
    uint64_t gen = 0;          // bumped whenever a buffer moves: captured graphs are stale
    bool dirty = true;         // bin counts / counters need a k_prep before the next frame
    uint32_t* host_counters = nullptr;   // pinned, CNT_COUNT words (then host_err, on its own cache line)
    uint32_t* host_err = nullptr;        // pinned word the kernels set nonzero with any error bit
    uint64_t* sig = nullptr;   // signal memory its stream writes after a framebuffer pass (null: use events)
};

}  // namespace axr

struct axr_ctx {
    int device = 0;
    uint64_t budget = 0;          // 0 = unlimited
    uint64_t in_use = 0;          // resource bytes (meshes, textures, framebuffers, batches, server pools)
    uint64_t arena_bytes = 0;     // all frame arenas of this context
    std::mutex mu;                // guards the two byte counters
    axr::ErrSink err;
    axr_config cfg{};
    axr::FrameArena main;
    // More arenas for asynchronous visibility-mode draws: consecutive such
    // draws take the arenas in turn (main is arena 0), each on its own
    // stream, so a frame's setup and shading run while earlier frames still
    // do.  A draw's framebuffer pass waits for everything queued on the
    // stream of the previous framebuffer work (which waited for its own
    // predecessor), and anything queued on main first waits for every other
    // stream (axr::ctx_join), so framebuffers see the calls in order.
    static constexpr int kArenas = 4;
    axr::FrameArena alt[kArenas - 1];
    hipEvent_t join_ev[kArenas] = {};   // join_ev[k]: recorded on arena k's stream for another to wait on
    hipEvent_t sync_ev[kArenas] = {};   // sync_ev[k]: recorded on arena k's stream for axr_sync to wait on
    int next_arena = 0;                 // arena of the next asynchronous visibility-mode draw
    int last_arena = 0;                 // arena whose stream took the latest framebuffer work
    bool alt_unjoined[kArenas - 1] = {};  // alt[k] took work main has not waited for
    axr::ArenaNeeds need_floor;         // the largest overflow needs reported: every arena is grown to them
    uint64_t last_value = 0;            // nonzero: that work was a draw whose stream then wrote this
                                        // value to its arena's signal (0: wait with an event)
    bool profiling = false;
    bool no_pass_skips = false;         // set if a pass judged empty on the host ever had work
    // timed frames not yet read back; the ring holds the frames of a whole
    // timed run, so timing does not stall the host until stats are read
    static constexpr int kEventRing = 1024;
    axr::StageEvents* ev = nullptr;     // kEventRing entries, created on first use
    int ev_head = 0, ev_pending = 0;
    axr_kernel_stats stats{};
    // axr_draw_noalloc: a call since the last sync may have left an overflow
    // in the device's scratch counters; pinned copy of those counters
    bool noalloc_pending = false;
    uint32_t* noalloc_counters = nullptr;
};

struct axr_texture {
    axr_ctx* ctx;
    axr::DevTexture dev;
    size_t bytes;
};

struct axr_mesh {
    axr_ctx* ctx;
    float* pos = nullptr;
    axr::Vertex56* verts = nullptr;
    uint32_t* idx = nullptr;
    uint32_t* face_colors = nullptr;
    axr::DevGroup* groups = nullptr;
    axr::DevMaterial* materials = nullptr;
    uint32_t ngroups = 0, nmaterials = 0;
    uint64_t nverts = 0, nfaces = 0;
    uint32_t nfaces_sub = 0;   // faces in submission order
    uint32_t batch_refs = 0;   // views of live batches that draw this mesh
    axr::DevGroup g0{};        // first group (host copy)
    size_t bytes = 0;
    // model-space bounds, for deciding on the host that a draw needs no
    // clipper or large-triangle pass: a sphere holding every vertex and the
    // longest edge of any face
    double center[3] = {0, 0, 0};
    double radius = 0, max_edge = 0;
};

struct axr_fb {
    axr_ctx* ctx;
    int w, h;
    uint32_t* color = nullptr;
    float* depth = nullptr;
    uint32_t* ids = nullptr;
    size_t bytes = 0;
    uint32_t batch_refs = 0;   // live batches with this framebuffer as a view
    // A clear not yet written: the next visibility-mode draw writes it as
    // part of its resolve; anything else that reads or draws into the planes
    // writes it first (axr::fb_settle).
    bool clear_pending = false;
    uint32_t clear_bgra = 0;
    float clear_depth = 0.0f;
};

namespace axr {

int fail(ErrSink& err, int code, const char* fmt, ...);
int hip_fail(ErrSink& err, hipError_t e, const char* what);
// Device allocation charged to the context's HBM budget.
int dev_alloc(axr_ctx* ctx, ErrSink& err, void** p, size_t bytes);
void dev_free(axr_ctx* ctx, void* p, size_t bytes);
int upload(ErrSink& err, hipStream_t st, void* dst, const void* src, size_t bytes);

// What one launch (a frame or a batch) needs from a frame arena.
struct FrameSize {
    uint64_t faces = 0;      // face slots over all views
    uint64_t tiles = 0;      // bins over all views
    uint64_t pixels = 0;     // visibility-buffer pixels (0 in binned mode)
    uint64_t color_slots = 0; // pixels of a two-pass resolve into a framebuffer with ids
    uint32_t wave_cap = 128; // pairs per setup-wave region (binned mode)
};
// High-water marks the kernels reported after an overflow.
FrameSize frame_size(const KArgs& ka, uint32_t wave_cap);

int arena_create(axr_ctx* ctx, ErrSink& err, FrameArena& fa);
void arena_destroy(axr_ctx* ctx, FrameArena& fa);
uint64_t arena_bytes(const FrameArena& fa);
int arena_ensure(axr_ctx* ctx, ErrSink& err, FrameArena& fa, const FrameSize& fs, const ArenaNeeds& need = {});
int arena_read_counters(ErrSink& err, FrameArena& fa);
// With the arena's work finished: the counters copied back if a kernel
// flagged an error in the pinned error word, else their host copy's error
// word zeroed.
int arena_counters_after_wait(ErrSink& err, FrameArena& fa);
// k_prep if the arena is fresh or dirty, then the frame's kernels; ev (may
// be null) times them.
int arena_launch(ErrSink& err, FrameArena& fa, const KArgs& ka, StageEvents* ev, const FbWait& wait = {});
// After arena_counters_after_wait (or arena_read_counters): AXR_OK, or grow the arena to the reported
// high-water marks, reset it and return AXR_ERR_OVERFLOW.
int arena_check_overflow(axr_ctx* ctx, ErrSink& err, FrameArena& fa, const FrameSize& fs);

// The uniforms drawMesh sets (reference tiled_pipeline.cpp:148-155).
// Queue on the main stream behind everything queued on the second arena's
// stream (no host wait).
int ctx_join(axr_ctx* ctx);
// Host wait for both streams.
int ctx_sync_all(axr_ctx* ctx);
// Write a framebuffer's pending clear, if any (stream-ordered).
int fb_settle(axr_fb* fb);
// After a host wait: the overflow an asynchronous axr_draw_noalloc left in
// the device's scratch counters (AXR_OK or AXR_ERR_OVERFLOW).
int noalloc_settle(axr_ctx* ctx);
// The shader kind all n views use, or -1 if they differ.
int uniform_shader(const ViewDesc* views, uint32_t n);
// Which of the passes after setup a draw of `m` with `mvp` on a W x H
// target can be shown to have no work (conservative: false when unsure).
struct PassSkips {
    bool clip = false;   // no face crosses a frustum plane
    bool big = false;    // and no triangle's pixel rectangle exceeds `footprint`
    // with clip: the back-face pre-test's constants (KArgs::cull_*)
    double eye[3] = {0.0, 0.0, 0.0};
    double sign = 0.0, tau = 0.0;
};
PassSkips pass_skips(const axr_mesh* m, const mat4& mvp, int W, int H, int footprint);
ViewDesc make_view(const axr_mesh* m, uint32_t* color, float* depth, uint32_t* ids, int w, int h,
                   const axr_draw_params* p, int tile);
bool shader_ok(const axr_draw_params* p);

}  // namespace axr
