// axr_types.hpp - plain data shared by the host side of the library and the
// gfx950 kernels: mesh vertices, materials, per-view draw descriptors, the
// triangle record, and the frame-arena layout.
#pragma once
#include <cstdint>
#include "axr_math.hpp"

namespace axr {

// Mesh vertex, 14 floats = 56 bytes, same field order as the reference's
// Vertex (reference: include/mesh.hpp:9-18).
struct alignas(8) Vertex56 {
    float px, py, pz;
    float u, v;
    float nx, ny, nz;
    float tx, ty, tz;
    float bx, by, bz;
};
static_assert(sizeof(Vertex56) == 56, "vertex layout");

// One RGBA8 texture resident in HBM: texel (x, y) is the uint32 at
// texels[y * w + x], red in the low byte.
struct DevTexture {
    const uint32_t* texels;
    int32_t w, h;
    int32_t filter;  // 0 nearest (the reference's sampling), 1 bilinear
    int32_t pad;
};

// Reference Material (include/mesh.hpp:20-34) as far as the shaders read it.
struct DevMaterial {
    DevTexture diffuse, metallic, bump, roughness, ao;
    float specular_exponent;
    float pad[3];
};

// One material group in submission order: faces [first, first + count) of
// the index buffer are submission ranks [rank_base, rank_base + count).
struct DevGroup {
    uint32_t first;
    uint32_t count;
    uint32_t rank_base;
    uint32_t material;
};

enum ShaderKind : int32_t {
    SHADER_FLAT = 0,
    SHADER_PHONG = 1,
    SHADER_PBR = 2,
    SHADER_GOURAUD = 3,
    SHADER_PHONG_POINT = 4,
    SHADER_FLAT_COLOR = 5,
    SHADER_COUNT = 6,
};

constexpr int MAX_POINT_LIGHTS = 4;

// Per-draw uniforms of the shaders (reference: include/IShader.hpp:39-45 plus
// the public fields of the shaders in include/shaders/shaders.hpp).
struct Uniforms {
    mat4 model;
    mat3 normal_mat;      // mat3(transpose(inverse(model))), for the flat shader
    vec3 camera_pos;
    vec3 light_dir;
    vec3 light_color;
    int32_t kind;
    int32_t num_lights;
    vec3 light_pos[MAX_POINT_LIGHTS];
    vec3 light_colors[MAX_POINT_LIGHTS];
};

// Everything the kernels need to know about one view (one drawMesh into one
// framebuffer).  A batch is an array of these; the plain path uses one.
struct ViewDesc {
    // mesh
    const float* pos;              // 3 floats per vertex (packed positions)
    const Vertex56* verts;         // full vertices, read only for shading winners
    const uint32_t* idx;           // 3 per face
    const uint32_t* face_colors;   // RGBA8 per face or null
    const DevGroup* groups;
    const DevMaterial* materials;
    const void* vary;              // the vertex stage's output for every vertex of the
                                   // mesh (the shader's Var type), computed by k_vary
                                   // earlier in the same launch; null: run it per pixel
    DevGroup g0;                   // groups[0], for single-group meshes without a load
    uint32_t ngroups;
    uint32_t nfaces;               // faces in submission order (sum of group counts)
    // transform
    mat4 mvp;
    Uniforms u;
    // framebuffer
    uint32_t* color;               // BGRA8 packed, row 0 = bottom
    float* depth;
    uint32_t* ids;                 // optional winner keys
    int32_t W, H;
    int32_t tiles_x, tiles_y;
    // offsets of this view inside the batch-wide arrays
    uint32_t face_base;            // first slot of this view in the record array
    uint32_t tile_base;            // first bin of this view
    uint32_t pix_base;             // first pixel of this view in the visibility buffer
    uint32_t inside;               // batches: the mesh's bounding sphere lies inside every
                                   // frustum plane, so k_setup leaves the classification out
};

// Triangle record, 48 bytes (three 16-byte loads).
//   a = (x0, y0, x1, y1)      screen position, y up
//   b = (x2, y2, z0, z1)      ndc z
//   c = (bits(z2), key, ref, info)
// key  = submission rank * 8 + clip sub-triangle: the depth tie-break order.
// For a record in the face-slot array, ref is the mesh face index; for a
// record in the clip array, ref is the view index.  A face slot whose face
// was clipped holds a redirect instead: info has REC_CLIPPED set and ref is
// the first clip record of that face (sub-triangle j sits at ref + j).
struct alignas(16) TriRec {
    float4 a;
    float4 b;
    uint4 c;
};
constexpr uint32_t REC_CLIPPED = 1u << 31;
constexpr uint32_t REC_MATERIAL_MASK = 0xFFFFu;
constexpr uint32_t REF_CLIP = 1u << 31;     // a bin entry with this bit names a clip record

// Size limits of one view (one drawMesh).  The submission key is
// rank * 8 + clip sub-triangle in 32 bits, so a view submits fewer than
// 2^29 faces; framebuffer edges are at most 32768 pixels.  Face slots of a
// whole batch are indexed in 31 bits (the top bit of a bin entry flags a
// clip record).
constexpr uint64_t MAX_VIEW_FACES = 1ull << 29;
constexpr int MAX_FB_EDGE = 32768;
constexpr uint64_t MAX_BATCH_FACES = 0x7FFFFFFFull;

// Counter block at the head of the frame arena.
enum Counter : uint32_t {
    CNT_CLIP_FACES = 0,   // faces queued for the full clipper
    CNT_CLIP_RECS = 1,    // clip records reserved
    CNT_OVF_PAIRS = 2,    // pairs placed in the overflow pool
    CNT_ACTIVE = 3,       // non-empty bins listed by the scan
    CNT_ERROR = 4,        // sticky error bits, see ERR_*
    CNT_NEED_CLIP_FACES = 5,
    CNT_NEED_CLIP_RECS = 6,
    CNT_NEED_PAIRS = 7,
    CNT_BIG_ITEMS = 8,    // visibility mode: work items of large triangles; with
    CNT_BIG_RECS = 9,     // the next word one 64-bit counter (records << 32 | items)
    CNT_NEED_BIG_RECS = 10,
    CNT_NEED_BIG_ITEMS = 11,
    CNT_COUNT = 16,
};
constexpr uint32_t ERR_CLIP_FACES = 1u;
constexpr uint32_t ERR_CLIP_RECS = 2u;
constexpr uint32_t ERR_PAIRS = 4u;
constexpr uint32_t ERR_BIG = 8u;
constexpr uint32_t ERR_SKIPPED = 16u;   // a pass the host skipped had work: the frame is redrawn with it

// Raster modes (axr_config.raster_mode).
enum RasterMode : int32_t {
    RASTER_BINNED = 0,       // setup -> bins -> per-bin LDS raster + resolve
    RASTER_VISIBILITY = 1,   // setup rasterizes small triangles into a per-pixel
                             // 64-bit key buffer; large ones go through work
                             // items; a per-bin pass resolves and shades
};

// Device frame arena: the per-frame working memory of the binning pipeline.
struct Arena {
    uint32_t* counters;        // CNT_COUNT words
    uint32_t* host_err;        // pinned host word set nonzero with every error bit, so a
                               // host wait learns of an error without copying the
                               // counters back (null: no such word)
    TriRec* rec;               // binned mode: one record per face slot of every view
    uint32_t* clip_base;       // visibility mode (same memory as rec): first clip
                               // record of each clipped face slot
    uint32_t* clip_faces;      // global face slots queued for clipping
    TriRec* clip_rec;          // clipped sub-triangles
    Vertex56* clip_verts;      // 3 per clip record
    uint32_t* tile_count;      // per bin
    uint32_t* tile_off;        // exclusive prefix of tile_count (+1 entry)
    uint4* pairs;              // (bin, slot in bin, record ref, 0) per overlap:
                               // wave_cap per setup wave, then the overflow pool
    uint32_t* wave_pairs;      // pairs each setup wave wrote into its own region
    uint32_t* bins;            // refs, grouped by bin
    uint32_t* active;          // non-empty bins, in bin order
    unsigned long long* scan_sums;   // per-scan-block partial sums (active << 32 | pairs)
    uint32_t cap_clip_faces;
    uint32_t cap_clip_recs;
    uint32_t cap_waves;        // setup waves with a region
    uint32_t wave_cap;         // pairs per setup-wave region (a wave that needs
                               // more takes them from the overflow pool with
                               // a single atomic)
    uint32_t cap_ovf;          // overflow pool entries (after the regions)
    uint32_t cap_bins;
    // visibility mode
    unsigned long long* vkeys; // per pixel of every view: (ordered depth << 32 | key),
                               // all ones when empty; the resolve leaves it empty
    uint32_t* vcol;            // two-pass resolve into a framebuffer with ids: each
                               // winner's colour, per pixel of every view
    TriRec* big_rec;           // large triangles (and clipped pieces) as records
    uint32_t* big_start;       // first work item of each large-triangle record
    uint32_t cap_vkeys;
    uint32_t cap_big_recs;
    uint32_t cap_big_items;
    uint32_t pad;
};

// One run of the vertex stage over a mesh's vertices for the views that
// share it (batched form): vertices [first, first + count) of the launch's
// numbering, the uniforms of view `view`, the shader's Var per vertex at out.
struct VaryJob {
    const Vertex56* verts;
    void* out;
    uint32_t first;
    uint32_t count;
    uint32_t view;
    int32_t kind;
};

// Kernel arguments common to every stage.  `table` points to `nviews`
// descriptors in device memory; when it is null the single descriptor `one`
// (passed by value) is the only view.
struct KArgs {
    const ViewDesc* table;
    const uint32_t* face_bases;   // nviews + 1 prefix (device), batch only
    const uint32_t* tile_bases;   // nviews + 1 prefix (device), batch only
    const uint32_t* tile_views;   // view of every bin (device), batch only
    uint32_t nviews;
    uint32_t total_faces;
    uint32_t total_tiles;
    int32_t tile_size;          // bin edge in pixels: 8, 16 or 32
    int32_t raster_threads;     // workgroup size of k_raster: 64, 128 or 256
    int32_t small_footprint;    // a triangle covering at most this many pixels
                                // of a bin is walked by one lane
    int32_t mode;               // RasterMode
    uint32_t total_pixels;
    int32_t shader = -1;        // the shader kind every view uses, -1 if they differ:
                                // selects a resolve kernel built for that shader alone
    int32_t clear = 0;          // visibility mode: k_resolve also clears every view's
    uint32_t clear_bgra = 0;    // framebuffer to these values (a clear the draw
    float clear_depth = 0.0f;   // absorbs instead of a pass of its own)
    int32_t two_pass = 0;       // plain visibility-mode draw with a clear: shade
                                // (k_shade) apart from the framebuffer writes
                                // (k_finish), which alone wait for earlier draws
    // batches: the vertex stage once per vertex of each shared mesh (k_vary)
    const VaryJob* vary_jobs = nullptr;
    const uint32_t* vary_firsts = nullptr;
    uint32_t vary_njobs = 0;
    uint32_t vary_total = 0;
    int32_t skip_clip = 0;      // one view, or every view of a binned batch: the mesh's
    int32_t skip_big = 0;       // bounding sphere lies inside every frustum plane, so
                                // no face is classified and the clipper is not
                                // launched (one view: nor, when no triangle can
                                // exceed the footprint, the large-triangle pass)
    // skip_clip views: k_vis's back-face pre-test in object space.  Every w
    // is then positive, so a face's screen-space winding has the sign of
    // cull_sign * ((p0 - cull_eye) . ((p1 - p0) x (p2 - p0))); a face whose
    // value lies below -max(cull_tau, 1e-3 |p0 - eye|_1 |p1 - p0|_1 |p2 - p0|_1)
    // covers more than 0.25 px^2 at the largest w of the mesh and is culled
    // by the exact screen-space test as well, so it is skipped before the
    // transform (cull_sign 0: no pre-test).
    float cull_eye[3] = {0.0f, 0.0f, 0.0f};
    float cull_sign = 0.0f;
    float cull_tau = 0.0f;
    int32_t depth_is_clear = 0; // binned mode: every framebuffer's depth plane holds
                                // clear_depth (written just before, by k_clear /
                                // k_clear_views): winners are tested against that
                                // value instead of a load of the plane
    ViewDesc one;
};

}  // namespace axr
