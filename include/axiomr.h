/* axiomr.h - C ABI of the MI355X tiled render path.
 *
 * The boundary sits where the reference renderer calls its pipeline:
 *   Pipeline::drawMesh(const glm::mat4& model, const Mesh&)
 *       reference: include/pipeline.hpp:25, src/tiled_pipeline.cpp:143-322
 * plus the resources drawMesh reads (Mesh, Material, Texture), the target it
 * writes (Framebuffer) and the camera uniforms it copies into the shader.
 * All signatures use plain pointers, sizes and POD structs; no exception
 * crosses the ABI.  Every call returns an axr_status (0 = ok); the message of
 * the last failure on a context is available from axr_last_error.
 *
 * Threading: a context is used by one host thread at a time.  Work is
 * enqueued on the context's HIP streams (four, one per frame arena;
 * asynchronous visibility-mode draws take them in turn and overlap, while
 * framebuffer writes keep call order); axr_draw is synchronous unless
 * AXR_DRAW_ASYNC is set, in which case completion (and any arena overflow) is
 * reported by axr_sync.  With AXR_OVERLAP=0 in the environment, or under a
 * profiler that collects hardware counters, draws do not overlap.
 */
#ifndef AXIOMR_H
#define AXIOMR_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define AXR_VERSION_MAJOR 0
#define AXR_VERSION_MINOR 1

typedef struct axr_ctx axr_ctx;
typedef struct axr_texture axr_texture;
typedef struct axr_mesh axr_mesh;
typedef struct axr_fb axr_fb;
typedef struct axr_batch axr_batch;
typedef struct axr_server axr_server;

typedef enum axr_status {
    AXR_OK = 0,
    AXR_ERR_INVALID = -1,    /* bad argument or handle */
    AXR_ERR_OOM = -2,        /* device or host allocation failed */
    AXR_ERR_HIP = -3,        /* a HIP runtime call failed */
    AXR_ERR_OVERFLOW = -4,   /* frame arena too small; frames since the last sync were not drawn */
    AXR_ERR_BUDGET = -5,     /* the request does not fit the HBM budget */
    AXR_ERR_NO_DEVICE = -6,  /* no usable GPU */
    AXR_ERR_NOT_READY = -7   /* a polled result is not complete yet */
} axr_status;

/* Shader kinds.  FLAT, PHONG and PBR restate the reference's FlatShader,
 * PhongShader and PBRShader (reference: include/shaders/shaders.hpp:19-62,
 * 140-247, 252-423).  GOURAUD, PHONG_POINT and FLAT_COLOR have no reference
 * counterpart: per-vertex Lambert times a texture, normal-mapped Phong with
 * point lights, and one RGBA8 colour per face. */
typedef enum axr_shader_kind {
    AXR_SHADER_FLAT = 0,
    AXR_SHADER_PHONG = 1,
    AXR_SHADER_PBR = 2,
    AXR_SHADER_GOURAUD = 3,
    AXR_SHADER_PHONG_POINT = 4,
    AXR_SHADER_FLAT_COLOR = 5
} axr_shader_kind;

typedef enum axr_filter {
    AXR_FILTER_NEAREST = 0,   /* reference Texture::sample (include/texture.hpp:12-33) */
    AXR_FILTER_BILINEAR = 1
} axr_filter;

/* reference Material (include/mesh.hpp:20-34): texture handles may be null. */
typedef struct axr_material {
    axr_texture* diffuse;
    axr_texture* metallic;
    axr_texture* bump;
    axr_texture* roughness;
    axr_texture* ao;
    float specular_exponent;
} axr_material;

/* reference MaterialGroup (include/mesh.hpp:35-39): faces
 * [first_face, first_face + face_count) drawn with materials[material].
 * Groups are drawn in array order. */
typedef struct axr_group {
    uint32_t material;
    uint32_t first_face;
    uint32_t face_count;
} axr_group;

#define AXR_DRAW_ASYNC 1u

/* Uniforms of one drawMesh call (reference tiled_pipeline.cpp:148-155 and the
 * shaders' public fields).  Matrices are column-major, 16 floats. */
typedef struct axr_draw_params {
    float model[16];
    float view_proj[16];
    float viewport[16];
    float camera_pos[3];
    int32_t shader;              /* axr_shader_kind */
    float light_dir[3];          /* FLAT, PHONG, PBR, GOURAUD */
    float light_color[3];        /* PHONG, PBR */
    int32_t num_lights;          /* PHONG_POINT: up to 4 */
    float light_pos[4][3];
    float light_colors[4][3];
    uint32_t flags;              /* AXR_DRAW_ASYNC */
} axr_draw_params;

/* Tunables of the binning pipeline (the autotuner writes these).  None of
 * them changes a pixel: coverage always follows the reference's 16-pixel
 * tiles. */
typedef struct axr_config {
    int32_t tile_size;           /* GPU bin edge: 8, 16 or 32 pixels (default 16) */
    int32_t raster_threads;      /* threads per bin workgroup: 64, 128 or 256 (default 64) */
    int32_t bin_capacity;        /* (bin, triangle) pairs each setup wave keeps in its own
                                    region, 64..1024 (default 128); more go to a shared pool */
    int32_t small_footprint;     /* triangles covering at most this many pixels of a bin are
                                    walked by one lane, 1..1024 (default 12) */
    int32_t raster_mode;         /* AXR_RASTER_BINNED (default) or AXR_RASTER_VISIBILITY */
    int32_t reserved[3];
} axr_config;

/* Raster modes; both give the same pixels.  BINNED: setup, per-bin lists
 * (prefix sum + scatter), one workgroup per bin rasterizes into LDS and
 * resolves.  VISIBILITY: setup rasterizes small triangles straight into a
 * per-pixel 64-bit (depth, submission) key buffer with atomics, large ones
 * as work items; a per-bin pass shades the winners.  tile_size is the
 * resolve bin there; bin_capacity is unused. */
enum { AXR_RASTER_BINNED = 0, AXR_RASTER_VISIBILITY = 1 };

/* Kernel time per stage, from HIP events attached to the stage kernels'
 * own dispatches, summed over the frames drawn while profiling is on
 * (frames keep overlapping).  Binned mode: 0 setup (transform,
 * classification, cull, triangle setup, bin counts), 1 clip, 2 prefix sum
 * of the bin counts, 3 bin scatter, 4 per-bin raster + resolve.
 * Visibility mode: 0 setup + small-triangle raster, 1 clip, 2 large-triangle
 * work items, 3 shading of a two-pass resolve (else unused), 4 resolve or
 * the two-pass framebuffer pass.  Both: 5 the batched form's vertex stage
 * per mesh vertex. */
typedef struct axr_kernel_stats {
    uint64_t frames;
    double frame_ms_total;
    double stage_ms_total[6];
    /* Counted whether or not profiling is on (reset with it): clipper and
     * large-triangle passes not launched because the draw's bounds showed
     * them empty, and draws in which such a pass turned out to have work
     * (redrawn with every pass; none are skipped on that context after). */
    uint64_t passes_skipped;
    uint64_t skip_faults;
} axr_kernel_stats;

/* ---- context ---------------------------------------------------------- */
const char* axr_version(void);
int axr_device_count(int* count);
int axr_ctx_create(int device, uint64_t hbm_budget_bytes, axr_ctx** out);
int axr_ctx_destroy(axr_ctx* ctx);
const char* axr_last_error(const axr_ctx* ctx);
int axr_ctx_set_config(axr_ctx* ctx, const axr_config* cfg);
int axr_ctx_get_config(const axr_ctx* ctx, axr_config* cfg);
int axr_ctx_set_profiling(axr_ctx* ctx, int enable);
int axr_ctx_get_stats(axr_ctx* ctx, axr_kernel_stats* out);
int axr_ctx_memory(const axr_ctx* ctx, uint64_t* bytes_in_use, uint64_t* arena_bytes);
int axr_sync(axr_ctx* ctx);

/* ---- resources (reference Texture, Mesh, Framebuffer) ----------------- */
/* rgba: w*h*4 bytes, row 0 first, as the reference stores stb_image output
 * (src/texture.cpp:21-35). */
int axr_texture_create(axr_ctx* ctx, const uint8_t* rgba, int w, int h, int filter, axr_texture** out);
int axr_texture_destroy(axr_texture* tex);

/* verts56: nverts x 14 floats laid out as the reference Vertex (position,
 * uv, normal, tangent, bitangent; include/mesh.hpp:9-18).  idx: 3 per face.
 * face_colors: optional RGBA8 per face (FLAT_COLOR shader), may be null. */
int axr_mesh_create(axr_ctx* ctx, const float* verts56, uint64_t nverts, const uint32_t* idx, uint64_t nfaces,
                    const axr_group* groups, uint32_t ngroups, const axr_material* materials, uint32_t nmaterials,
                    const uint32_t* face_colors, axr_mesh** out);
int axr_mesh_destroy(axr_mesh* mesh);

/* BGRA8 colour + FP32 depth (reference Framebuffer, include/framebuffer.hpp).
 * with_ids != 0 adds a per-pixel winner-key plane (debugging/parity). */
int axr_fb_create(axr_ctx* ctx, int w, int h, int with_ids, axr_fb** out);
int axr_fb_destroy(axr_fb* fb);
/* reference Framebuffer::clearColor/clearDepth (src/framebuffer.cpp:26-42):
 * bgra packed as a<<24 | r<<16 | g<<8 | b.  Not synchronous: the clear is
 * recorded, and the next axr_draw in visibility mode writes it in the same
 * pass that writes the frame (no pass of its own); a read-back, a binned
 * draw or a batch writes it first.  The pixels are always as if the clear
 * had run at this call. */
int axr_fb_clear(axr_fb* fb, uint32_t bgra, float depth);
/* Synchronous read-back; any pointer may be null. */
int axr_fb_read(axr_fb* fb, uint8_t* bgra, float* depth, uint32_t* ids);

/* ---- the hot path ------------------------------------------------------ */
/* reference TiledPipeline::drawMesh (src/tiled_pipeline.cpp:143-322). */
int axr_draw(axr_ctx* ctx, axr_fb* fb, const axr_mesh* mesh, const axr_draw_params* params);

/* axr_draw without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory while
 * they draw (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  The same visibility-mode kernels and pixels as
 * axr_draw, on the context's own stream, with the frame's working memory
 * carved out of one fixed 64 MiB device array per GPU (held by the process
 * from library load until it exits) instead of the context's frame arenas.
 * A frame that needs more than 64 MiB (about 1920x1080 with 1M faces needs
 * 30 MB; the per-pixel keys take 8 bytes a pixel) fails with
 * AXR_ERR_INVALID before any work is queued; a frame whose clipped or
 * large triangles overflow their share is not drawn and reports
 * AXR_ERR_OVERFLOW (at the call, or at axr_sync with AXR_DRAW_ASYNC).  All
 * calls on one GPU share its array: calls on one context are ordered by its
 * stream, calls from different contexts or threads that overlap race. */
int axr_draw_noalloc(axr_ctx* ctx, axr_fb* fb, const axr_mesh* mesh, const axr_draw_params* params);

/* Size limits of one view (one drawMesh), checked by axr_draw, batch
 * creation and server submission: AXR_OK, or AXR_ERR_INVALID when the view
 * submits 2^29 faces or more (the depth tie-break key, submission rank * 8
 * + clip sub-triangle, is 32 bits; reference tiled_pipeline.cpp:249-258
 * orders by the same (group, face, sub-triangle) triple) or the
 * framebuffer is not 1..32768 pixels on each edge.  Host only: no device
 * is needed. */
int axr_check_view(uint64_t submitted_faces, int width, int height);

/* ---- batched form: many views in one graph-captured launch sequence ---- */
/* Each view i draws meshes[i] into fbs[i] with params[i].  Views have no
 * order among themselves, so each framebuffer may appear in one view only
 * (AXR_ERR_INVALID otherwise); several meshes into one framebuffer, as the
 * reference's Renderer::render loop draws them (src/renderer.cpp:100-110),
 * are consecutive axr_draw calls or one batch per mesh.  The batch holds
 * its views' framebuffers and meshes: axr_fb_destroy and axr_mesh_destroy
 * fail with AXR_ERR_INVALID on them until the batch is destroyed. */
int axr_batch_create(axr_ctx* ctx, uint32_t nviews, axr_fb* const* fbs, const axr_mesh* const* meshes,
                     const axr_draw_params* params, axr_batch** out);
/* Clear every view's framebuffer (bgra packed as in axr_fb_clear, depth)
 * at the start of each launch, inside the same graph. */
int axr_batch_set_clear(axr_batch* batch, int enable, uint32_t bgra, float depth);
int axr_batch_launch(axr_batch* batch);    /* enqueue (graph replay); completes at axr_sync */
int axr_batch_destroy(axr_batch* batch);

/* ---- served form: continuous batching of render requests --------------- */
/* No counterpart in the reference (it renders one window's frames).  Each
 * request is one drawMesh (src/tiled_pipeline.cpp:143-322) of a resident
 * mesh into a framebuffer of the requested size, cleared to clear_bgra and
 * +inf depth first (src/framebuffer.cpp:6-14).  Requests are admitted in
 * arrival order as lanes and HBM free up, and drawn in batches on `lanes`
 * streams, each with its own frame arena, so batches overlap on the device.
 * A request's pixels do not depend on what else shares its batch.  Meshes,
 * textures and the context must outlive the requests that use them. */
typedef struct axr_server_config {
    int lanes;                  /* batches in flight at once, 1..8 (0: 3) */
    uint32_t max_batch_views;   /* requests per batch launch (0: 64) */
    uint64_t max_batch_faces;   /* submitted triangles per batch; a larger
                                   single request runs alone (0: 8 Mi) */
    uint64_t max_batch_tiles;   /* bins per batch, and per request (0: 1 Mi) */
    uint64_t hbm_budget;        /* bytes for frame arenas + framebuffers;
                                   0 = no limit beyond the context's */
    uint32_t clear_bgra;        /* clear colour of every request */
    int with_ids;               /* keep each request's winner-key plane */
    uint64_t max_batch_pixels;  /* framebuffer pixels per batch; a larger single
                                   request runs alone (0: 32 Mi).  Each lane's
                                   frame arena is sized for max_batch_faces
                                   faces over this many pixels up front */
} axr_server_config;

typedef struct axr_serve_request {
    const axr_mesh* mesh;
    int width, height;
    axr_draw_params params;     /* flags ignored */
} axr_serve_request;

typedef enum axr_request_state {
    AXR_REQ_QUEUED = 1,         /* waiting for a lane or for HBM */
    AXR_REQ_RUNNING = 2,
    AXR_REQ_DONE = 3,
    AXR_REQ_FAILED = 4
} axr_request_state;

typedef struct axr_server_stats {
    uint64_t submitted, completed, failed, rejected;
    uint64_t batches, batched_views, max_views_in_batch;
    uint64_t launches_overlapped;   /* batches launched while another lane ran */
    uint32_t max_lanes_busy;
    uint32_t reserved0;
    uint64_t deferred;              /* admission passes stopped by the budget */
    uint64_t fb_reserved_bytes, fb_peak_bytes, fb_in_use_bytes;
    uint64_t fixed_bytes;           /* frame arenas and batch tables */
    uint64_t queue_peak;
    uint64_t overflow_retries;      /* batches redrawn after an arena overflow */
} axr_server_stats;

/* Starts the scheduler thread.  The context's tunables (axr_config) at this
 * call are the server's.  Fails with AXR_ERR_BUDGET when the lanes' frame
 * arenas alone exceed hbm_budget. */
int axr_server_create(axr_ctx* ctx, const axr_server_config* cfg, axr_server** out);
/* Drains running batches, then frees everything (unreleased results too). */
int axr_server_destroy(axr_server* server);
const char* axr_server_last_error(axr_server* server);
/* Queues one request.  AXR_ERR_BUDGET if its framebuffer can never fit the
 * budget next to the arenas; a request that fits later waits in the queue. */
int axr_server_submit(axr_server* server, const axr_serve_request* req, uint64_t* ticket);
/* *state = axr_request_state; returns the request's error once FAILED. */
int axr_server_poll(axr_server* server, uint64_t ticket, int* state);
/* AXR_OK when done, AXR_ERR_NOT_READY on timeout, else the request's error. */
int axr_server_wait(axr_server* server, uint64_t ticket, double timeout_s);
/* Copies a done request's planes to the host (any pointer may be null). */
int axr_server_read(axr_server* server, uint64_t ticket, uint8_t* bgra, float* depth, uint32_t* ids);
/* Returns the request's framebuffer to the pool; the ticket becomes invalid. */
int axr_server_release(axr_server* server, uint64_t ticket);
/* pause != 0: requests keep queueing but none is admitted; 0: admit the
 * queue in order again (a burst submitted while paused is batched). */
int axr_server_pause(axr_server* server, int pause);
int axr_server_get_stats(axr_server* server, axr_server_stats* out);
/* Open-loop driver: request i is submitted at (call time + arrival_s[i]),
 * non-decreasing; results are released as they complete and latency_s[i]
 * receives completion minus scheduled arrival in seconds (-1 if failed).
 * Returns after the last request completes. */
int axr_server_replay(axr_server* server, uint32_t n, const axr_serve_request* reqs, const double* arrival_s,
                      double* latency_s);

/* ---- ingestion and image output (host side) ---------------------------- */
/* reference Mesh(const std::string&) (src/mesh.cpp:8-27, 305-415): a
 * Wavefront OBJ file, its materials read from <dir>/<basename>.mtl
 * (src/mesh.cpp:180-215; mtllib is ignored, as there), vertices
 * deduplicated on (position, uv, normal), polygons fanned, tangents and
 * bitangents generated per vertex (src/mesh.cpp:222-303).  Faces before the
 * first usemtl belong to no group. */
typedef struct axr_obj axr_obj;
typedef struct axr_obj_group {
    char material[128];
    uint64_t first_face, face_count;
    uint32_t material_index;      /* into axr_obj_get_material order; UINT32_MAX if the
                                     .mtl has no such material (reference: .at() throws) */
} axr_obj_group;
typedef struct axr_obj_material {  /* reference Material (include/mesh.hpp:20-34) */
    char name[128];
    float ambient[3], diffuse[3], specular[3];
    float specular_exponent, refractive_index, dissolve;
    int32_t illumination_model;
    /* texture paths (directory of the .mtl + the map_* argument), "" if none:
     * map_Kd; map_Ks or refl; map_Bump, bump or norm; map_Ns; map_A0 */
    char diffuse_map[512], metallic_map[512], bump_map[512], roughness_map[512], ao_map[512];
} axr_obj_material;
int axr_obj_load(const char* path, axr_obj** out);
int axr_obj_destroy(axr_obj* obj);
const char* axr_obj_last_error(void);              /* this thread's last failure */
int axr_obj_counts(const axr_obj* obj, uint64_t* nverts, uint64_t* nfaces, uint32_t* ngroups,
                   uint32_t* nmaterials);
/* nverts x 14 floats (the 56-byte vertex of axr_mesh_create), nfaces x 3 */
int axr_obj_arrays(const axr_obj* obj, float* verts56, uint32_t* idx);
int axr_obj_get_group(const axr_obj* obj, uint32_t i, axr_obj_group* out);
int axr_obj_get_material(const axr_obj* obj, uint32_t i, axr_obj_material* out);  /* sorted by name */

/* reference Texture(const std::string&) (src/texture.cpp:21-35): the decoded
 * RGBA8 image, top row first (stbi_load with 4 channels).  PNG (all colour
 * types, 1-16 bit, tRNS, Adam7), JPEG (baseline and progressive, the bytes
 * stb_image gives) and binary PPM/PGM. */
int axr_image_load(const char* path, int* w, int* h, uint8_t** rgba);  /* free with axr_image_free */
void axr_image_free(uint8_t* rgba);
const char* axr_image_last_error(void);            /* this thread's last failure */
int axr_image_write_png(const char* path, int w, int h, const uint8_t* rgba, int flip_rows);
int axr_texture_load(axr_ctx* ctx, const char* path, int filter, axr_texture** out);
/* The framebuffer's colour as an RGBA PNG, top row first. */
int axr_fb_write_png(axr_fb* fb, const char* path);

/* ---- synthetic scenes (host only; no device work) ---------------------- */
/* UV sphere of radius r with `stacks` x `slices` quads: (stacks+1)*(slices+1)
 * vertices, 2*stacks*slices faces. */
int axr_gen_uv_sphere(int stacks, int slices, float radius, float* verts56, uint32_t* idx);
/* Geodesic icosphere, frequency n: 10n^2+2 vertices, 20n^2 faces, radius
 * 1 + amp * valueNoise4(p) (4 octaves), lattice hashed with `seed`. */
int axr_gen_icosphere(int n, float amp, uint64_t seed, float* verts56, uint32_t* idx);
/* Screen-space triangle soup for a camera at the origin looking down -z
 * (fov_deg vertical, aspect W/H): centres uniform over the W x H screen,
 * screen area log-normal (median_area pixels, sigma), one view depth per
 * triangle uniform in [0.1, 100], counter-clockwise (front-facing) corners.
 * A dup_frac share of the triangles repeats an earlier triangle's corners
 * exactly (with its own colour) to force exact depth ties.  Writes 3*ntri
 * vertices, ntri faces and ntri RGBA8 colours. */
int axr_gen_soup(uint64_t ntri, int W, int H, float fov_deg, float median_area, float sigma, float dup_frac,
                 uint64_t seed, float* verts56, uint32_t* idx, uint32_t* colors);
int axr_gen_checker(int w, int h, int cell, uint8_t* rgba);
int axr_gen_noise_texture(int w, int h, uint64_t seed, uint8_t* rgba);
int axr_gen_normal_map(int w, int h, float min_z, uint64_t seed, uint8_t* rgba);
/* reference Camera (src/camera.cpp:30-81): view, projection (RH, z in
 * [-w, w]) and projection*view, column-major; plus the viewport matrix
 * (camera.cpp:191-205).  The matrices of a camera constructed at pos
 * looking at target, after its first update(). */
int axr_camera_matrices(const float pos[3], const float target[3], float fov_deg, float aspect, float znear,
                        float zfar, int vp_w, int vp_h, float* view16, float* proj16, float* view_proj16,
                        float* viewport16);

/* ---- fly camera (reference src/camera.cpp, include/camera.hpp) ---------
 * The reference's Camera reads a Window each frame (handleInput) and then
 * advances (update(dt)).  Here the window's answers arrive as an explicit
 * axr_camera_input, so a recorded input sequence replays to the same
 * camera path on every run; axr_camera_path produces a whole path's
 * matrices in one call, ready to be drawn as one batch.  The state record
 * is caller-owned (no allocation); matrices are column-major. */

/* Movement keys held during a frame (reference camera.cpp:116-124). */
enum { AXR_KEY_W = 1, AXR_KEY_S = 2, AXR_KEY_A = 4, AXR_KEY_D = 8 };

/* What Camera::handleInput asks the Window (camera.cpp:100-127): the mouse
 * position, whether the right button (button 1) is down, the keys held. */
typedef struct axr_camera_input {
    float mouse_x, mouse_y;
    int32_t right_button;
    uint32_t keys;                  /* AXR_KEY_* bits */
} axr_camera_input;

/* The reference Camera's members (camera.hpp:38-72). */
typedef struct axr_camera {
    float position[3];
    float orientation[4];           /* quaternion w, x, y, z */
    float forward[3], right[3], up[3];
    float delta_yaw, delta_pitch;   /* mouse deltas times sensitivity, from the last input */
    float velocity[3];              /* sum of the held keys' directions, unnormalized */
    float fov, aspect, znear, zfar;
    float mouse_sensitivity;
    float last_mouse[2];
    int32_t viewport[4];            /* x, y, width, height */
    float view[16], proj[16], view_proj[16];
    int32_t view_dirty, proj_dirty;
} axr_camera;

/* Camera::Camera (camera.cpp:30-55): fov clamped to [1, 179], aspect to
 * >= 0.1; the projection is built, the view stays identity until the first
 * axr_camera_update, as in the reference. */
int axr_camera_init(axr_camera* cam, const float position[3], const float target[3], float fov_deg, float aspect);
/* Camera::handleInput (camera.cpp:100-127). */
int axr_camera_handle_input(axr_camera* cam, const axr_camera_input* input);
/* Camera::update (camera.cpp:83-98): rotates only while both mouse deltas
 * are non-zero, moves by velocity*dt, rebuilds the view if it changed. */
int axr_camera_update(axr_camera* cam, float dt);
/* Camera setters (camera.cpp:162-189).  As in the reference, a new fov or
 * aspect marks the projection stale but no later call rebuilds it, so the
 * projection keeps the constructor's fov and aspect. */
int axr_camera_set_position(axr_camera* cam, const float position[3]);
int axr_camera_set_target(axr_camera* cam, const float target[3]);
int axr_camera_set_fov(axr_camera* cam, float fov_deg);
int axr_camera_set_aspect_ratio(axr_camera* cam, float aspect);
int axr_camera_set_viewport(axr_camera* cam, int x, int y, int width, int height);
/* Camera::getViewportMatrix (camera.cpp:191-205). */
int axr_camera_viewport_matrix(const axr_camera* cam, float* viewport16);
/* n frames of the reference's frame loop (src/renderer.cpp:122-123:
 * handleInput, then update(dt)), writing each frame's projection*view
 * (16 floats, may be NULL) and position (3 floats, may be NULL). */
int axr_camera_path(axr_camera* cam, const axr_camera_input* inputs, const float* dts, uint64_t n,
                    float* view_proj16n, float* positions3n);

#ifdef __cplusplus
}
#endif

#endif /* AXIOMR_H */
